/*
 * pmkd.h -- C ABI of the MI355X kd-tree build + ray traversal library
 * (libpmkd.so).
 *
 * Plain pointers and sizes only.  Device pointers are HIP device memory on
 * the context's device.  Every "stream" argument is a hipStream_t passed as
 * void* (NULL = the null stream).  Every function returns PM_OK (0) or an
 * error code; pm_last_error() returns the message of the last failure on
 * that context.
 *
 * The reference (OpenEngineDK extensions-PhotonMapping) has no C ABI: its
 * path is a set of C++ classes.  Each entry point below names the reference
 * interface it stands in for (file:line, relative to the reference root).
 * The C++ classes with the reference's names live in include/pmkd/ (one header per class) and
 * call only these functions.
 */
#ifndef PMKD_H
#define PMKD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PMKD_ABI_VERSION 3

/* error codes */
#define PM_OK                  0
#define PM_ERR_ARG             1  /* invalid argument                       */
#define PM_ERR_NOT_IMPLEMENTED 2  /* TriangleMapUpperCreator.cu:551-552      */
#define PM_ERR_DEPTH           3  /* upper build exceeded max_upper_levels   */
#define PM_ERR_INTERNAL        4  /* TriangleMapUpperCreator.cu:606-607      */
#define PM_ERR_UNSUPPORTED     5  /* option not provided by this library     */
#define PM_ERR_HIP             6  /* HIP runtime error (CHECK_FOR_CUDA_ERROR) */
#define PM_ERR_NO_TREE         7  /* trace before build / set_tree           */
#define PM_ERR_CAPACITY        8  /* a captured build outgrew its capacities */

/* TriangleMap::LowerAlgorithm, Utils/CUDA/TriangleMap.h:33 */
#define PM_LOWER_BITMAP   0
#define PM_LOWER_BALANCED 1
#define PM_LOWER_SAH      2
/* TriangleMap::SplitMethod, Utils/CUDA/TriangleMap.h:34; SPLIT is not
 * implemented (PM_ERR_NOT_IMPLEMENTED, as TriangleMapUpperCreator.cu:551) */
#define PM_SPLIT_BOX    0
#define PM_SPLIT_DIVIDE 1
#define PM_SPLIT_SPLIT  2
/* IRayTracer::IntersectionAlgorithm, IRayTracer.h:30 */
#define PM_INTERSECT_MOELLER 0
#define PM_INTERSECT_WOOP    1
#define PM_RAY_ORDER_INPUT    0
#define PM_RAY_ORDER_COHERENT 1
/* tracer kinds: BruteTracer, RayTracer (kd-restart), ShortStack */
#define PM_TRACE_BRUTE       0
#define PM_TRACE_KD_RESTART  1
#define PM_TRACE_SHORT_STACK 2

typedef struct pm_ctx pm_ctx;

/* Hit record (16 bytes): triangle id (index into p0/p1/p2, -1 on a miss),
 * distance t and barycentrics u (weight of p1) and v (weight of p2), the
 * values MoellerTrumbore leaves in (primHit, tHit) (IRayTracer.h:111-145).
 * A miss is {-1, +inf, 0, 0}. */
typedef struct { int id; float t, u, v; } pm_hit;

/* Build options: the TriangleMap / creator setters.
 *   lower_algorithm       TriangleMap::SetLowerAlgorithm     TriangleMap.cu:68-81
 *   empty_space_splitting TriangleMap::SplitEmptySpace       TriangleMap.cu:83-85
 *   empty_space_threshold SetEmptySpaceThreshold             TriangleMapUpperCreator.h:86
 *   split_method          TriangleMap::SetSplitMethod        TriangleMap.cu:90
 *   propagate_aabbs       TriangleMap::SetPropagateBoundingBox TriangleMap.cu:96-102
 *   traversal_cost        TriangleMap::SetTraversalCost      TriangleMap.cu:104-108
 *   lower_size            KDNODE_MAX_LOWER_SIZE              KDNode.h:21 (<= 64)
 *   max_upper_levels      no counterpart: depth cap (0 = 128); the reference
 *                         loops forever on >= lower_size coincident triangles
 *   sah_leaf_max          no counterpart: force SAH splits above this many
 *                         triangles (0 = off)
 *   sah_empty_bonus       no counterpart: cost factor for a split with one
 *                         empty side (1 = off)
 *   max_ref_factor        no counterpart: error (PM_ERR_DEPTH) when a level
 *                         holds more than max_ref_factor * n + 65536
 *                         references (0 = 16); the reference doubles its
 *                         references every level on coincident triangles
 *   chunk_size            refs per upper-stage segment, the reference's
 *                         Segments::SEGMENT_SIZE (Segments.h:27, 256);
 *                         128/256/512/1024/2048/4096, 0 = 1024 (the
 *                         fastest on MI355X for a 1M-triangle build).
 *                         Changes how the work is cut, never the tree
 *                         (pm_autotune picks it) */
typedef struct {
    int   lower_algorithm;
    int   empty_space_splitting;
    float empty_space_threshold;
    int   split_method;
    int   propagate_aabbs;
    float traversal_cost;
    int   lower_size;
    int   max_upper_levels;
    int   sah_leaf_max;
    float sah_empty_bonus;
    int   max_ref_factor;
    int   chunk_size;
} pm_build_params;

/* Trace options: IRayTracer setters (IRayTracer.h:50-58) and constants.
 *   leaf_skipping  IRayTracer::SetLeafSkipping (default on, IRayTracer.cu:33-34);
 *                  forced off when the tree has no propagated boxes
 *                  (Raytracer.cu:313)
 *   stack_size     SHORT_STACK_SIZE (ShortStack.cu:25), 1..16
 *   brute_cap      MAX_PRIMS (BruteTracer.cu:50); 0 = test every triangle
 *   intersection   IRayTracer::SetIntersectionAlgorithm (IRayTracer.h:52):
 *                  PM_INTERSECT_MOELLER (default) or PM_INTERSECT_WOOP, the
 *                  latter on per-triangle Woop matrices computed at each
 *                  call (GeometryList::GetWoopValues, GeometryList.cu:103-117)
 *   ray_order      no reference counterpart: PM_RAY_ORDER_INPUT (default)
 *                  traces the rays in input order; PM_RAY_ORDER_COHERENT
 *                  lets the tree tracers run a trace of >= 131072 rays in
 *                  the order of a spatial key (origin cell, direction cell;
 *                  mesh first in a forest) so a wave's rays walk similar
 *                  paths -- for incoherent rays (photons, diffuse bounces,
 *                  random probes); rays that are coherent already (camera
 *                  rays in pixel order) only pay for the sort.  Each ray's
 *                  record (and counters) is the same either way and goes to
 *                  the ray's own index; brute force and the canvas keep the
 *                  input order.
 *   step_cap       no reference counterpart: 0 (default) or the most inner
 *                  nodes a ray's descent walks per round of the tree
 *                  tracers' wave-cooperative leaf test before it rejoins
 *                  the wave (a longer descent continues in the next round),
 *                  so rays with short descents do not wait for the wave's
 *                  longest one -- for descent-heavy rays (grazing camera
 *                  rays over terrain); records and counters are the same
 *                  for every value.  0 <= step_cap <= 4096. */
typedef struct {
    int leaf_skipping;
    int stack_size;
    int brute_cap;
    int intersection;
    int ray_order;
    int step_cap;
} pm_trace_params;

/* Tree view: device pointers into the context's tree (non-owning; valid
 * until the next build/set_tree/destroy), in the reference's SoA layout
 * (KDNode.h:65-73, TriangleNode.h:20-21):
 *   info[n]             0 LEAF, 1 X, 2 Y, 3 Z
 *   split_pos[n]
 *   aabb_min/max[4n]    propagated node boxes (float4)
 *   primitive_index[n]  first slot in prim_indices
 *   primitive_bitmap[n] leaves: mask over prim_indices[primitive_index + i];
 *                       inner upper nodes: the reference count
 *   children[2n]        (left, right); left is the side below split_pos
 *   surface_area[n], parent[n]
 *   prim_indices[m]     triangle ids (index into p0/p1/p2)
 *   leaf_ids[k]         upper-stage leaves in creation order
 * num_roots: the roots are nodes 0..num_roots-1, one per mesh of a
 * batched build (1 for pm_build; 0 is read as 1 by pm_set_tree). */
typedef struct {
    int num_nodes;
    const int8_t *info;
    const float *split_pos;
    const float *aabb_min;
    const float *aabb_max;
    const int *primitive_index;
    const long long *primitive_bitmap;
    const int *children;
    const float *surface_area;
    const int *parent;
    int num_prim_indices;
    const int *prim_indices;
    int num_leaf_ids;
    const int *leaf_ids;
    const float *p0, *p1, *p2;
    int num_tris;
    int upper_levels, lower_levels;
    int has_boxes;
    int num_roots;
} pm_tree;

/* Tree statistics (PrintTree, TriangleMap.cu:110-116, plus depth/leaf data). */
typedef struct {
    int num_nodes, reachable_nodes, inner_nodes, leaves, empty_leaves;
    int max_depth;
    double mean_leaf_depth;
    long long leaf_refs;
    int max_leaf_size;
    int upper_levels, lower_levels;
    int num_prim_indices, num_upper_leaves;
    /* work counts of the last build (no reference counterpart): refs of every
     * upper level summed, refs those levels wrote (children + leaves), and
     * prims of every node the SAH lower levels examined */
    long long upper_ref_reads, upper_ref_writes, lower_prims;
    /* 1 when the last build ran with capacities derived from the build before
     * it in this context (one read-back at the end), 0 when it was sized level
     * by level (a read-back per level); the tree is the same either way */
    int planned_build;
} pm_stats;

/* Defaults of the TriangleMap constructor (TriangleMap.cu:27-49) and the
 * upper creator (TriangleMapUpperCreator.cu:28-32): BITMAP lower creator,
 * empty-space splitting on at 0.35, BOX split, boxes propagated, traversal
 * cost 24, lower size 32. */
void pm_build_params_default(pm_build_params *p);
/* IRayTracer constructor defaults (IRayTracer.cu:32-38), ShortStack size 4,
 * no brute-force cap. */
void pm_trace_params_default(pm_trace_params *p);

/* Context = TriangleMap + the three tracers' device state.  Replaces
 * TriangleMap::TriangleMap (TriangleMap.cu:27-49) and the tracer
 * constructors (Raytracer.cu:38-53, ShortStack.cu:37-52, BruteTracer.cu:30-46). */
int pm_create(pm_ctx **out, int device);
int pm_destroy(pm_ctx *ctx);
const char *pm_last_error(const pm_ctx *ctx);

/* TriangleMap::Create (TriangleMap.cu:51-62): upper stage
 * (TriangleMapUpperCreator::Create, .cu:110-183) then the lower creator
 * (ITriangleMapCreator::Create, ITriangleMapCreator.h:45-46).  p0/p1/p2 are
 * device float4[n] (the GeometryList SoA, GeometryList.h:58-68); they are
 * borrowed and must stay valid while the tree is traced.  Stream-ordered;
 * the call returns after the build's last host read-back. */
int pm_build(pm_ctx *ctx, const float *p0, const float *p1, const float *p2, int n,
             const pm_build_params *params, void *stream);

/* Geometry only, for the brute-force tracer (BruteTracer(GeometryList*),
 * BruteTracer.h:28). */
int pm_set_geometry(pm_ctx *ctx, const float *p0, const float *p1, const float *p2, int n);

/* TriangleMap::GetNodes / GetPrimitiveIndices (TriangleMap.h:77-79) and the
 * KDNode getters (KDNode.h:85-94). */
int pm_tree_view(pm_ctx *ctx, pm_tree *out);

/* Upload a tree given in HOST memory (same layout as pm_tree, host
 * pointers; p0/p1/p2 in the struct are DEVICE pointers).  Lets a tree built
 * elsewhere be traced. */
int pm_set_tree(pm_ctx *ctx, const pm_tree *host_tree, void *stream);

/* TriangleMap::PrintTree (TriangleMap.cu:110-116) as numbers: walks the tree
 * on the host (copies it back). */
int pm_tree_stats(pm_ctx *ctx, pm_stats *out, void *stream);

/* IRayTracer::Trace (IRayTracer.h:47; RayTracer::Trace Raytracer.cu:295-390,
 * ShortStack::Trace ShortStack.cu:231-321, BruteTracer::Trace
 * BruteTracer.cu:123-172) with the ray arrays passed in (float4 org/dir per
 * ray, device) and the first hit of each ray written to out (device) instead
 * of a shaded canvas.  counters (device int[4*nrays] or NULL): per ray
 * inner-node visits, leaf visits, triangle tests, leaf-box tests. */
int pm_trace(pm_ctx *ctx, int algo, const float *org, const float *dir, int nrays,
             pm_hit *out, const pm_trace_params *params, int *counters, void *stream);

/* No counterpart in the reference: pm_trace for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  Same arguments, records and counters as pm_trace;
 * allocates nothing: the work counter and the coherent-order sort buffers are
 * a file-scope device array of the library (about 17 MB, held on every GPU the
 * library runs on until the process ends).  Limits: at most
 * PM_NOALLOC_MAX_ORDERED_RAYS rays when the coherent order applies (more:
 * PM_ERR_ARG; use PM_RAY_ORDER_INPUT, which has no limit), MOELLER
 * intersection only (WOOP: PM_ERR_UNSUPPORTED).  The tree (pm_build /
 * pm_set_tree) must exist.  Every call on one GPU shares that GPU's array:
 * calls that overlap on one GPU (two streams, two threads) race. */
#define PM_NOALLOC_MAX_ORDERED_RAYS (1 << 20)
int pm_trace_noalloc(pm_ctx *ctx, int algo, const float *org, const float *dir, int nrays,
                     pm_hit *out, const pm_trace_params *params, int *counters, void *stream);

/* Batched build: one kd-tree per mesh for M meshes whose triangles are
 * concatenated in p0/p1/p2 (mesh m owns triangles
 * [mesh_offsets[m], mesh_offsets[m+1]); mesh_offsets is HOST int[M+1],
 * non-decreasing, from 0 to n).  All M trees are built together, level by
 * level, into one node array (a forest): root m is node m, and each tree is
 * the one pm_build makes for its mesh alone, up to a renumbering of nodes and
 * prim slots that keeps their order (hit ids are ids into the concatenated
 * arrays).  Replaces a loop of TriangleMap::Create (TriangleMap.cu:51-62)
 * over the meshes of a GeometryList (GeometryList.h:40-68), which in the
 * reference builds one tree over all of them. */
int pm_build_batched(pm_ctx *ctx, const float *p0, const float *p1, const float *p2, int n,
                     const int *mesh_offsets, int num_meshes, const pm_build_params *params,
                     void *stream);

/* pm_trace over a forest: ray r starts at root ray_mesh[r] (device int[nrays];
 * a ray whose mesh index is outside [0, num_meshes) misses).  KD_RESTART and
 * SHORT_STACK only.  Same traversal as IRayTracer::Trace (see pm_trace). */
int pm_trace_batched(pm_ctx *ctx, int algo, const float *org, const float *dir, const int *ray_mesh,
                     int nrays, pm_hit *out, const pm_trace_params *params, int *counters,
                     void *stream);

/* Capacities of a captured batch (pm_batch_graph_capture).  In: 0 fields
 * are derived from an eager build of the same batch (1.5x what it reached
 * + 1024; levels + 2 upper, + 1 lower); out: what was captured. */
typedef struct {
    long long nodes;          /* tree nodes */
    long long active;         /* active nodes of one upper level */
    long long segments;       /* chunks of one upper level */
    long long level_refs;     /* references of one upper level */
    long long leaf_refs;      /* leaf slots (prim indices) */
    long long leaf_ids;       /* upper-stage leaves */
    long long lower_nodes;    /* nodes of one lower (SAH) level */
    int upper_levels, lower_levels;
} pm_graph_caps;

/* One graph launch per frame (no reference counterpart; the reference
 * rebuilds and traces with host-driven launches, PhotonRenderingView.cpp:
 * 34-85).  Records the batched build (as pm_build_batched) followed by the
 * batched trace (as pm_trace_batched, algo KD_RESTART or SHORT_STACK) into
 * a hipGraph owned by the context: every level's sizes live in device
 * memory, buffers are allocated at the capacities beforehand, and no host
 * read-back or synchronisation happens inside.  All pointers are baked in:
 * each replay rebuilds the trees from the CURRENT contents of p0/p1/p2
 * (e.g. after pm_deform_radial + pm_add_mesh) and traces org/dir into out;
 * the mesh offsets, sizes and parameters are fixed.  Capture runs one eager
 * build first (for the capacities) and one checked replay.  A replay whose
 * geometry outgrows the capacities stops early with PM_ERR_CAPACITY
 * reported by pm_batch_graph_status; capture again then. */
int pm_batch_graph_capture(pm_ctx *ctx, const float *p0, const float *p1, const float *p2, int n,
                           const int *mesh_offsets, int num_meshes, const float *org, const float *dir,
                           const int *ray_mesh, int nrays, pm_hit *out, int algo,
                           const pm_build_params *build_params, const pm_trace_params *trace_params,
                           pm_graph_caps *caps, void *stream);
/* Replay: one hipGraphLaunch on `stream`, asynchronous. */
int pm_batch_graph_launch(pm_ctx *ctx, void *stream);
/* Wait for the last replay and report its build status (PM_OK or the error
 * that stopped it); also makes the tree's sizes visible to pm_tree_view /
 * pm_tree_stats (which call it themselves). */
int pm_batch_graph_status(pm_ctx *ctx, int *build_status, void *stream);

/* Autotuner result (pm_autotune). */
typedef struct {
    int chunk_size, lower_size, stack_size;   /* the chosen setting */
    float build_ms, trace_ms;                 /* its median times */
    int trials;                               /* settings timed */
    int ray_order;                            /* the chosen pm_trace_params.ray_order */
    int step_cap;                             /* the chosen pm_trace_params.step_cap */
} pm_tuning;

/* Autotuner (no reference counterpart: the reference fixes these at
 * compile time, Segments.h:27 SEGMENT_SIZE, KDNode.h:21 lower size,
 * ShortStack.cu:25 stack size).  Times the caller's own workload on this
 * GPU: chunk_size over {256, 512, 1024, 2048} (build time), then lower_size
 * (the upper stage's leaf threshold) over {8, 16, 24, 32, 48, 64} (cost =
 * builds_per_trace * build + trace), then stack_size over {4, 8, 16}
 * together with ray_order over {INPUT, COHERENT} and step_cap over {0, 4}
 * (trace time); each setting is timed `reps` times (median) after one
 * warm-up; a leaf threshold whose build stops at the reference-count
 * guard (PM_ERR_DEPTH) is skipped.  Every other field comes from `base`.  The geometry is one mesh
 * (mesh_offsets NULL) or a batch (as pm_build_batched); the rays are traced
 * with SHORT_STACK (ray_mesh as pm_trace_batched, NULL for one mesh).  On
 * return the context holds the tree of the chosen setting. */
int pm_autotune(pm_ctx *ctx, const float *p0, const float *p1, const float *p2, int n,
                const int *mesh_offsets, int num_meshes, const float *org, const float *dir,
                const int *ray_mesh, int nrays, const pm_build_params *base,
                float builds_per_trace, int reps, pm_tuning *out, void *stream);

/* Light of the canvas shading: ColorKernels.h:10-13 (__constant__
 * d_lightPosition / Ambient / Diffuse / Specular). */
typedef struct {
    float position[3], ambient[3], diffuse[3], specular[3];
} pm_light;

/* IRayTracer::Trace(IRenderCanvas*, uchar4*) (IRayTracer.h:47) with the
 * rays passed in: per ray, the tracer's first hit is shaded with
 * PhongLighting and the ray turned into its mirror reflection (Lighting,
 * ColorKernels.h:15-77), colours are blended (BlendColor, :79-84) and the
 * reflected ray is traced again, until a miss, alpha >= 0.97 or max_bounces
 * (the reference has no cap); canvas gets one uchar4 per ray (device).
 * n0/n1/n2: device float4 corner normals per triangle, c0: device uchar4
 * colour per triangle (GeometryList's n0..n2/c0, see pm_add_mesh_attributes).
 * Shadows are off, as in the reference (Raytracer.cu:219-226).  The
 * specular power is the correctly rounded float of the double-precision
 * power (the reference's device __powf is not reproduced). */
int pm_trace_canvas(pm_ctx *ctx, int algo, const float *org, const float *dir, int nrays,
                    const float *n0, const float *n1, const float *n2, const unsigned char *c0,
                    const pm_light *light, int max_bounces, unsigned char *canvas,
                    const pm_trace_params *params, void *stream);

/* The normal and colour half of AddMeshKernel (float4/uchar4 variant,
 * GeometryList.cu:197-232): corner normals (normalMat * n.xyz, 0) with
 * normalMat the upper-left 3x3 of the row-major model matrix (identity when
 * NULL), and corner colours, for triangles [offset, offset + nidx/3).
 * normals: device float4[nv]; colors: device uchar4[nv] or NULL (then c0..c2
 * are left alone). */
int pm_add_mesh_attributes(pm_ctx *ctx, const float *normals, const unsigned char *colors, int nv,
                           const unsigned *indices, int nidx, const float *model, float *n0, float *n1,
                           float *n2, unsigned char *c0, unsigned char *c1, unsigned char *c2, int offset,
                           void *stream);

/* Backward of the intersection (no reference counterpart; the forward is
 * MoellerTrumbore, IRayTracer.h:111-145).  p0/p1/p2 (device float4[ntris])
 * are the triangles the forward traced (the geometry of the build that made
 * the hits); the context's own geometry is neither used nor changed.  For
 * every ray with hits[r].id >= 0 (the records a pm_trace call produced for
 * these rays) and upstream gradients grad_t/grad_u/grad_v (device
 * float[nrays] each, NULL = zero): grad_org/grad_dir (device float[3*nrays],
 * optional) get the per-ray gradients; grad_vertices (device
 * float[3*num_vertices]) gets the per-vertex sums over corners, where faces
 * (device int[3*ntris]) maps triangle corners to vertex ids (NULL: vertex
 * id = 3*triangle + corner, so num_vertices must be >= 3*ntris).  The sums
 * are deterministic: entries are added in (ray, corner) order after a stable
 * radix sort by vertex id.  A hit id outside [-1, ntris), or a face index of
 * a hit triangle outside [0, num_vertices), gives PM_ERR_ARG (the outputs
 * are then unspecified; nothing is read or written out of range).
 * Synchronises the stream (one 4-byte read-back at the end). */
int pm_trace_backward(pm_ctx *ctx, const float *p0, const float *p1, const float *p2, int ntris,
                      const float *org, const float *dir, const pm_hit *hits, int nrays,
                      const float *grad_t, const float *grad_u, const float *grad_v, const int *faces,
                      int num_vertices, float *grad_vertices, float *grad_org, float *grad_dir,
                      void *stream);

/* GeometryList::AddMesh(CUDAMeshNode*, modelMat) (GeometryList.cu:318-341,
 * AddMeshKernel :197-232): gather an indexed mesh (device float4 vertices[nv],
 * device uint indices[nidx], nidx a multiple of 3) into the triangle SoA,
 * writing p0/p1/p2[offset .. offset + nidx/3) (device float4).  model is a
 * HOST row-major 4x4 matrix (NULL = identity) applied as p = M v with each
 * row dotted left to right; a caller holding an OpenEngine model matrix
 * passes its transpose, as the reference does.  Out-of-range indices give
 * PM_ERR_ARG.  Synchronises the stream. */
int pm_add_mesh(pm_ctx *ctx, const float *vertices, int nv, const unsigned *indices, int nidx,
                const float *model, float *p0, float *p1, float *p2, int offset, void *stream);

/* OscCUDAMeshNode::Handle / UpdateSurface (OscCUDAMeshNode.cu:72-107): one
 * animation step of the oscillating grid.  vertices (device float4[d*d],
 * d = detail + 1, vertex (x, y) at y*d + x as OscCUDAMeshNode::Init lays it
 * out) get y = height + amplitude cos(frequency x + time)
 * + amplitude sin(frequency y + time); normals (device float4[d*d] or NULL)
 * the matching unnormalised normal.  The reference uses amplitude 0.2,
 * frequency 0.3 and time = accumulated milliseconds / 700000. */
int pm_osc_surface(pm_ctx *ctx, float *vertices, float *normals, int detail, float height,
                   float time, float amplitude, float frequency, void *stream);

/* Per-frame deformer of unit-sphere meshes (the animated spheres a batched
 * rebuild traces; no reference counterpart, written in the style of
 * UpdateSurface): out[i] = unit[i] * r with
 * r = 1 + wave[0] sin(wave[1] theta + wave[2] frame) cos(wave[3] phi + wave[4] frame),
 * theta = acos(z), phi = atan2(y, x); w is copied.  unit/out: device float4[nv]. */
int pm_deform_radial(pm_ctx *ctx, const float *unit, int nv, float frame, const float *wave,
                     float *out, void *stream);

/* IRayTracer::CreateInitialRays / CreateRays / Unproject (IRayTracer.cu:
 * 45-100): one camera ray per pixel of a width x height canvas, ray
 * id = y * width + x, written to org/dir (device float4[width * height]).
 * inv_view_proj (HOST, 16 floats) is the inverse of view * projection in
 * the element order Unproject reads (the screen point (s.x, s.y, 1, 1) maps
 * to c_k = m[k] s.x + m[4+k] s.y + m[8+k] + m[12+k]); cam_pos (HOST, 3
 * floats) is the camera position.  s = ((x + jx) / width) * 2 - 1, likewise
 * for y: jitter 0 gives jx = jy = 0 (the reference), jitter 1 a seeded
 * sub-pixel offset in [0, 1)^2 from PCG32 stream (jitter_seed,
 * 1 << 40 | id).  Direction = normalize(c.xyz / c.w - cam_pos), w = 0;
 * origin = (cam_pos, 1).  The caller computes the inverse (the reference
 * calls OpenEngine's Matrix::GetInverse). */
int pm_create_rays(pm_ctx *ctx, const float *inv_view_proj, const float *cam_pos, int width, int height,
                   unsigned long long jitter_seed, int jitter, float *org, float *dir, void *stream);

/* Largest light triangle count pm_photon_emit takes. */
#define PM_MAX_LIGHTS 65536

/* Photon emission (no reference counterpart; the caller of the photon
 * trace): n photons from the area lights l0/l1/l2 (device float4[num_lights],
 * 1 <= num_lights <= PM_MAX_LIGHTS, e.g. a slice of p0/p1/p2).  Photon i
 * uses PCG32 stream (seed, 5 << 40 | i): it picks light k with probability
 * proportional to its area (the first k with u0 * total < running area sum,
 * all sums FP32 in light order), a uniform point p of that triangle and a
 * cosine-weighted direction about n = normalize((v1 - v0) x (v2 - v0)), the
 * side the light emits to; origin = p + offset * n (w = 1), direction w = 0,
 * into out_org/out_dir (device float4[n]).  Lights of zero total area give
 * undefined rays. */
int pm_photon_emit(pm_ctx *ctx, const float *l0, const float *l1, const float *l2, int num_lights, int n,
                   unsigned long long seed, float offset, float *out_org, float *out_dir, void *stream);

/* Photon path continuation (no reference counterpart: the reference's only
 * secondary rays are the shading loop's reflections, Raytracer.cu:219-256).
 * For every ray i with hits[i].id in [0, absorb_from): one diffuse
 * continuation ray, origin o + t d + offset * n (n the geometric normal turned
 * against d), direction cosine-weighted about n from PCG32 stream
 * (seed, 6 << 40 | i).  Survivors are compacted in input order into
 * out_org/out_dir (device float4[n]); out_src[j] (device, optional) is the
 * source index; *n_out (host) receives the count, and out_count (device,
 * optional) too.  Uses the geometry of the last build/set_geometry. */
int pm_photon_bounce(pm_ctx *ctx, const float *org, const float *dir, const pm_hit *hits, int n,
                     int absorb_from, unsigned long long seed, float offset, float *out_org,
                     float *out_dir, int *out_src, int *out_count, int *n_out, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* PMKD_H */
