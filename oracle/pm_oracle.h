/*
 * pm_oracle.h -- CPU oracle for the kd-tree build + ray traversal path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: the tests, the
 * smoke check and the CPU-baseline leg of the benchmark link it; the product
 * (libpmkd) never does.
 *
 * It restates, in plain C++ with IEEE single precision and no FMA
 * contraction, the semantics of the reference's kd-tree construction and
 * traversal (OpenEngineDK extensions-PhotonMapping):
 *   traversal  Utils/CUDA/IRayTracer.h:60-168, Raytracer.cu:57-262,
 *              ShortStack.h:27-98, ShortStack.cu:58-194, BruteTracer.cu:54-96
 *   upper      Utils/CUDA/TriangleMapUpperCreator.cu:110-759 and
 *              Kernels/{TriangleUpper,TriangleUpperSegment,ReduceSegments,
 *              TriangleUpperChildren,EmptySpaceSplitting}.h
 *   lower      Utils/CUDA/TriangleMapSAHCreator.cu:53-246,
 *              Kernels/LowerTriangleMap.h:24-232,
 *              TriangleMapBitmapCreator.cu:32-65
 * Host (#else) branches are followed: IEEE division, no __fdividef.
 */
#ifndef PM_ORACLE_H
#define PM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_tree orc_tree;

/* Build options; field meaning matches pm_build_params in include/pmkd.h. */
typedef struct {
    int   lower_algorithm;       /* 0 BITMAP, 2 SAH */
    int   empty_space_splitting; /* TriangleMapUpperCreator.cu:31 default 1 */
    float empty_space_threshold; /* TriangleMapUpperCreator.cu:28 0.35 */
    int   split_method;          /* 0 BOX (only one supported) */
    int   propagate_aabbs;       /* TriangleMap.cu:47 default 1 */
    float traversal_cost;        /* TriangleMap.cu:48 24 */
    int   lower_size;            /* KDNode.h:21 32 (64 allowed) */
    int   max_upper_levels;      /* depth cap (no counterpart); 0 = 128 */
    int   sah_leaf_max;          /* 0 = off */
    float sah_empty_bonus;       /* 1 = off */
    int   max_ref_factor;        /* refs guard: error above f*N + 65536 (0 = 16) */
} orc_build_params;

typedef struct {
    int leaf_skipping;  /* IRayTracer.cu:33-34 default 1 */
    int stack_size;     /* ShortStack.cu:25 default 4 */
    int brute_cap;      /* BruteTracer.cu:50 4096 in the reference; 0 = all */
    int threads;        /* OpenMP threads, 0 = runtime default */
    int intersection;   /* IRayTracer::IntersectionAlgorithm (IRayTracer.h:30): 0 MOELLER, 1 WOOP */
} orc_trace_params;

typedef struct { int id; float t, u, v; } orc_hit;

typedef struct {
    int num_nodes, reachable_nodes, inner_nodes, leaves, empty_leaves;
    int max_depth;
    double mean_leaf_depth;      /* over non-empty leaves */
    long long leaf_refs;         /* sum of popcount over reachable leaves */
    int max_leaf_size;
    int upper_levels, lower_levels;
    int num_prim_indices, num_upper_leaves;
} orc_stats;

/* build; returns 0 on success, non-zero with orc_last_error() set. */
int  orc_build(const float *p0, const float *p1, const float *p2, int n,
               const orc_build_params *params, orc_tree **out);
void orc_free(orc_tree *t);
const char *orc_last_error(void);
/* forest of one tree per mesh (mesh m = triangles [mesh_offsets[m],
 * mesh_offsets[m+1])): the upper stage starts from num_meshes roots, nodes
 * 0..num_meshes-1, processed together level by level; orc_build is the
 * one-mesh case. */
int  orc_build_batched(const float *p0, const float *p1, const float *p2, int n,
                       const int *mesh_offsets, int num_meshes,
                       const orc_build_params *params, orc_tree **out);
int  orc_num_roots(const orc_tree *t);
/* WoopTransformationMatrix(a, b, c) -> m[12] = m0, m1, m2 (float4 each) */
void orc_woop_matrix(const float *a, const float *b, const float *c, float *m);
/* the balanced lower creator's decision for one node (CalcSplit,
 * LowerTriangleMap.h:299-369): returns split (0/1), outputs axis, position,
 * child masks and the winning small^2 + large^2 */
int  orc_balanced_node(int n, const float *bmin, const float *bmax, long long bitmap,
                       float traversal_cost, int lower_size, int *axis, float *pos,
                       long long *left, long long *right, int *cost);
/* TriangleAabbIntersectionStep3 (the DIVIDE split's triangle/box test) */
int  orc_tri_box_step3(const float *v0, const float *v1, const float *v2,
                       const float *bmin, const float *bmax);

/* tree arrays (copies are owned by the tree) */
int  orc_num_nodes(const orc_tree *t);
int  orc_num_prim_indices(const orc_tree *t);
int  orc_num_leaf_ids(const orc_tree *t);
int  orc_upper_levels(const orc_tree *t);
/* refs entering upper level l (l < orc_upper_levels) */
long long orc_upper_level_refs(const orc_tree *t, int level);
long long orc_lower_level_nodes(const orc_tree *t, int level);
int  orc_lower_levels(const orc_tree *t);
void orc_copy_tree(const orc_tree *t, int8_t *info, float *split_pos,
                   float *aabb_min /*4N*/, float *aabb_max /*4N*/,
                   int *primitive_index, long long *primitive_bitmap,
                   int *children /*2N*/, float *surface_area, int *parent,
                   int *prim_indices, int *leaf_ids);
int  orc_stats_of(const orc_tree *t, orc_stats *s);

/* Build a tree object from caller-supplied arrays (for tracing a tree that
 * came from somewhere else, e.g. the GPU builder). */
int  orc_tree_from_arrays(int num_nodes, const int8_t *info,
                          const float *split_pos, const float *aabb_min,
                          const float *aabb_max, const int *primitive_index,
                          const long long *primitive_bitmap,
                          const int *children, int num_prim_indices,
                          const int *prim_indices, orc_tree **out);

/* traversal; algo 0 brute, 1 kd-restart, 2 short-stack.
 * org/dir are float4 per ray (w ignored).  counters (optional, int[4] per
 * ray): inner-node visits, leaf visits, triangle tests, leaf-box tests. */
int  orc_trace(const orc_tree *t, const float *p0, const float *p1,
               const float *p2, int ntris, int algo, const float *org,
               const float *dir, int nrays, const orc_trace_params *params,
               orc_hit *out, int *counters);
/* The tracers' canvas output (IRayTracer::Trace(canvas), IRayTracer.h:47):
 * per ray, first hit -> PhongLighting + mirror reflection -> BlendColor,
 * until a miss, alpha >= 0.97 or max_bounces; canvas = uchar4 per ray.
 * n0/n1/n2 float4 corner normals and c0 uchar4 colour per triangle
 * (GeometryList, GeometryList.cu:197-232); light = position, ambient,
 * diffuse, specular (3 floats each). */
int  orc_trace_canvas(const orc_tree *t, const float *p0, const float *p1, const float *p2,
                      int ntris, int algo, const float *org, const float *dir, int nrays,
                      const float *n0, const float *n1, const float *n2,
                      const unsigned char *c0, const float *light, int max_bounces,
                      const orc_trace_params *params, unsigned char *canvas);
/* the same over a forest: ray r starts at root ray_root[r] (a ray whose
 * root is not one of the forest's misses); kd-restart and short-stack. */
int  orc_trace_batched(const orc_tree *t, const float *p0, const float *p1,
                       const float *p2, int ntris, int algo, const float *org,
                       const float *dir, const int *ray_root, int nrays,
                       const orc_trace_params *params, orc_hit *out, int *counters);

/* single Moller-Trumbore test exactly as the tracers run it (FP32), and the
 * same formula evaluated in FP64 (no acceptance test, raw t/u/v/det). */
int  orc_mt32(const float *v0, const float *v1, const float *v2,
              const float *o, const float *d, float t_max, float *tuv);
void orc_mt64(const double *v0, const double *v1, const double *v2,
              const double *o, const double *d, double *tuvdet);

/* The SAH decision for ONE node, exposed for brute-force checks:
 * n prims with boxes bmin/bmax (3 floats each, slot order) and areas,
 * bitmap of the node.  Outputs: split flag, axis (1..3), split position,
 * left/right masks, left/right areas, cost, node area. */
int  orc_sah_node(int n, const float *bmin, const float *bmax,
                  const float *areas, long long bitmap, float node_area,
                  float traversal_cost, int sah_leaf_max,
                  float sah_empty_bonus, int *axis, float *pos,
                  long long *left, long long *right, float *left_area,
                  float *right_area, float *cost);

/* triangle area 0.5*|e1 x e2| as LowerTriangleMap.h:32-36 computes it */
float orc_tri_area(const float *v0, const float *v1, const float *v2);

/* FP64 intersection adjoint: for a fixed triangle, given dL/dt, dL/du,
 * dL/dv, returns dL/dv0, dL/dv1, dL/dv2 (3 each), dL/do, dL/dd (3 each). */
void orc_mt64_adjoint(const double *v0, const double *v1, const double *v2,
                      const double *o, const double *d, double gt,
                      double gu, double gv, double *g_v0, double *g_v1,
                      double *g_v2, double *g_o, double *g_d);

/* Backward of the intersection over a ray batch (see pm_trace_backward in
 * include/pmkd.h).  FP32 restates the product's operation order exactly and
 * sums per-vertex entries sequentially in (ray, corner) order; FP64 uses
 * orc_mt64_adjoint and the same summation order in double.  faces NULL =
 * vertex id 3*triangle + corner.  Outputs are overwritten (grad_vertices
 * zeroed first); grad_org/grad_dir optional. */
int orc_trace_backward32(const float *p0, const float *p1, const float *p2, int ntris,
                         const float *org, const float *dir, const orc_hit *hits, int nrays,
                         const float *grad_t, const float *grad_u, const float *grad_v,
                         const int *faces, int num_vertices, float *grad_vertices,
                         float *grad_org, float *grad_dir);
int orc_trace_backward64(const float *p0, const float *p1, const float *p2, int ntris,
                         const float *org, const float *dir, const orc_hit *hits, int nrays,
                         const float *grad_t, const float *grad_u, const float *grad_v,
                         const int *faces, int num_vertices, double *grad_vertices,
                         double *grad_org, double *grad_dir);

/* Camera rays exactly as pm_create_rays makes them (IRayTracer.cu:45-70):
 * m = inverse view-projection in Unproject's element order, cam = camera
 * position, jitter 0/1 with PCG32 stream (seed, 1 << 40 | id); org/dir
 * float[4 * width * height]. */
void orc_create_rays(const float *m, const float *cam, int width, int height, unsigned long long seed,
                     int jitter, float *org, float *dir);

#ifdef __cplusplus
}
#endif
#endif
