// pm_internal.h -- host-side internals of libpmkd: device buffers, the
// context (tree + scratch), and the entry points of the build and trace
// translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "../../include/pmkd.h"

namespace pmkd {

// Error carrying a pmkd status code.
struct Error {
    int code;
    std::string msg;
};

#define PM_HIP(call)                                                                    \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess)                                                           \
            throw ::pmkd::Error{PM_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)}; \
    } while (0)

#define PM_LAUNCH_CHECK() PM_HIP(hipGetLastError())

// Grow-only device allocation (the CUDADataBlock<T> role: Extend never
// shrinks, contents are kept when it grows).
template <class T>
struct DevBuf {
    T *ptr = nullptr;
    size_t cap = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { if (ptr) (void)hipFree(ptr); }
    // ensure capacity >= n elements; keeps the first `keep` elements
    void reserve(size_t n, hipStream_t s, size_t keep = 0) {
        if (n <= cap) return;
        size_t ncap = cap ? cap : 256;
        while (ncap < n) ncap += ncap / 2 + 256;
        T *np = nullptr;
        PM_HIP(hipMalloc(&np, ncap * sizeof(T)));
        if (ptr && keep) PM_HIP(hipMemcpyAsync(np, ptr, keep * sizeof(T), hipMemcpyDeviceToDevice, s));
        if (ptr) {
            PM_HIP(hipStreamSynchronize(s));
            PM_HIP(hipFree(ptr));
        }
        ptr = np;
        cap = ncap;
    }
};

// Pinned host staging for small read-backs.
struct HostPinned {
    int *ptr = nullptr;
    size_t n = 0;
    ~HostPinned() { if (ptr) (void)hipHostFree(ptr); }
    void ensure(size_t count) {
        if (count <= n) return;
        if (ptr) PM_HIP(hipHostFree(ptr));
        PM_HIP(hipHostMalloc((void **)&ptr, count * sizeof(int), hipHostMallocDefault));
        n = count;
    }
};

// Node arrays in the reference's SoA layout.
struct NodeArrays {
    DevBuf<int8_t> info;
    DevBuf<float> split;
    DevBuf<float4> bmin, bmax;
    DevBuf<int> prim_index;
    DevBuf<long long> bitmap;
    DevBuf<int2> children;
    DevBuf<float> area;
    DevBuf<int> parent;
    void reserve(size_t n, hipStream_t s, size_t keep) {
        info.reserve(n, s, keep); split.reserve(n, s, keep);
        bmin.reserve(n, s, keep); bmax.reserve(n, s, keep);
        prim_index.reserve(n, s, keep); bitmap.reserve(n, s, keep);
        children.reserve(n, s, keep); area.reserve(n, s, keep); parent.reserve(n, s, keep);
    }
};

// Device view used by the tracers.
struct TreeDev {
    const int8_t *info;
    const float *split;
    const int2 *children;
    const int *prim_index;
    const long long *bitmap;
    const float4 *bmin, *bmax;
    const int *prim_indices;
    const int4 *packed;        // one 16-byte record per node (trace view)
    const float4 *leaf_tris;   // 3 float4 per leaf slot: v0 (w = id bits), e1, e2
    const float4 *leaf_woop;   // Woop intersection: 3 float4 per leaf slot (m0, m1, m2)
    int num_nodes;
    int num_roots;             // roots are nodes 0..num_roots-1 (one per mesh)
    long long num_slots;       // leaf slots (3 float4 each in leaf_tris)
    const int *fail;           // device flag: nonzero = the tree is not valid, every ray misses
};

// Level state of a build, in device memory: the plan kernels write it
// between the phases of a level, the level kernels read their sizes from
// it.  The eager build reads it back to size buffers and grids; a captured
// build never does (caps bound every write, a failed check sets status).
struct Lvl {
    int A, active_index, nodes_size, R, S, E;          // upper level
    int L_all, R_all, LL_all, LR_all, leaf_nodes;      // this level's counts
    int leafTri, nextR, created, child_start, next_active;
    int final_refs, nleaf, level;
    int An, list_base, use_list, splits, lower_base, lower_levels;   // lower levels
    int cand_n;                                        // lower level: nodes listed for the full SAH search
    int advance;                                       // the children plan is stored, the level not yet advanced
    // a level's active nodes, their first index and segments as the level's
    // classify and leaf passes read them, by level parity: written before
    // the level starts (the level before's leaf pass, or the build's start)
    // and not changed while those passes run
    int scanA[2], scanAI[2], scanS[2];
    int scanR[2];                                      // the level's box keys are replicated (few nodes before)
    int SL_all;                                        // next-level segments of this level's left children
    int leaf_left;                                     // leaf slots among the level's left children
    // work counts of the build (the algorithmic-bytes formula's inputs)
    unsigned long long ref_reads;                      // refs of every upper level, summed
    unsigned long long ref_writes;                     // refs written by every upper level (children + leaves)
    unsigned long long lower_prims;                    // prims of every node of every SAH lower level
    // the largest sizes the build reached (the capacities of a later build derive from them)
    long long seen_active, seen_segs, seen_refs, seen_lower;
    int status;                                        // 0, or the PM_ERR_* that stopped the build
    int max_levels;
    long long max_refs;
    long long cap_nodes, cap_active, cap_segs, cap_refs, cap_leaf_refs, cap_leaf_ids, cap_lower;
};

// A copy of the level state in coherent pinned host memory: an eager build's
// upper levels (the last tile of the leaf pass) and lower levels (the plan
// kernel) write it there, each word tagged with the level's sequence number;
// the host waits for a record whose words all carry the number instead of
// copying the state back and synchronising the stream.
struct LvlMirror {
    static constexpr int NW = (int)(sizeof(Lvl) / sizeof(int));
    // word t of the level state in the low half of w[t], the number of the
    // level that wrote it in the high half: the host takes the record once
    // every word carries the number it waits for, so the device needs no
    // ordering between the words' stores
    struct Rec {
        unsigned long long w[NW];
    };
    Rec *host = nullptr;   // coherent, mapped
    Rec *dev = nullptr;    // the device's address of it
    unsigned next_seq = 0;
    // the next level's number: never 0, the tag of a record no level wrote
    unsigned take_seq() {
        if (++next_seq == 0u) next_seq = 1u;
        return next_seq;
    }
    ~LvlMirror() { if (host) (void)hipHostFree(host); }
    bool failed = false;   // no such memory here: the build copies the state back instead
    // true when the mirror is usable; PMKD_HOST_MIRROR=0 in the environment
    // (read when a context's first eager build starts) makes that context's
    // builds copy the state back instead
    bool ensure() {
        if (host || failed) return !failed;
        const char *env = std::getenv("PMKD_HOST_MIRROR");
        if (env && env[0] == '0') {
            failed = true;
            return false;
        }
        if (hipHostMalloc((void **)&host, sizeof(Rec), hipHostMallocCoherent | hipHostMallocMapped) != hipSuccess ||
            hipHostGetDevicePointer((void **)&dev, host, 0) != hipSuccess) {
            (void)hipGetLastError();
            if (host) (void)hipHostFree(host);
            host = nullptr;
            dev = nullptr;
            failed = true;
            return false;
        }
        for (int t = 0; t < NW; ++t) __atomic_store_n(&host->w[t], 0ull, __ATOMIC_RELEASE);
        return true;
    }
};

// Sizes an eager build reached (the capacities a captured build derives
// from) and the fixed capacities / level counts of a captured build.
struct BuildSizes {
    long long nodes = 0, active = 0, segs = 0, refs = 0, leaf_refs = 0, leaf_ids = 0, lower = 0;
    int upper_levels = 0, lower_levels = 0;
};

// Build scratch (kept across builds; grow-only).
struct BuildScratch {
    DevBuf<float4> ref_min[2], ref_max[2];     // level ref arrays (ping-pong)
    DevBuf<float4> prim_min, prim_max;         // final (upper leaf) refs
    DevBuf<int> keys[2];                        // 6 ordered-int box keys per active node
    DevBuf<int> node_segbase[2];                // per active node: first segment (by level parity)
    DevBuf<int> seg_node[2], seg_start[2], seg_count[2];   // segment tables (by level parity)
    DevBuf<int> segL, segR;                     // per segment: left / right refs
    DevBuf<int> segLoff, segRoff;               // per segment: left / right refs before it
    DevBuf<int2> node_lna, node_lref;           // per active node: leaf slots and leaf refs before it (left, right)
    DevBuf<int2> node_lseg;                     // per active node: its non-leaf children's segments before it
    DevBuf<int4> node_sz;                       // per active node: children sizes, side offsets of its first segment
    DevBuf<unsigned long long> lb_words[2];     // look-back status words of the leaf passes (by level parity)
    DevBuf<int> lb_ctr;                         // their tile tickets (by level parity)
    DevBuf<int> scan_tmp[3];                    // block sums for scans, per level
    DevBuf<int> totals;                         // small device-side counters
    DevBuf<unsigned long long> prims_part;      // lower stage: prims counted per wave slot (64)
    DevBuf<float> areas;                        // lower stage: triangle areas per final ref
    DevBuf<int> lower_list;                     // lower stage: nodes listed for the full SAH search
    DevBuf<int> lower_flag, lower_off;
    DevBuf<long long> lower_sets;               // 2 per node
    DevBuf<float> lower_careas;                 // 2 per node
    DevBuf<int> bounce_flags, bounce_offs;      // photon driver compaction
    DevBuf<float> light_cdf;                    // photon emission: running sum of the light areas
    DevBuf<unsigned> bwd_keys[2], bwd_vals[2];  // backward: radix sort ping-pong
    DevBuf<float> bwd_corner;                   // backward: 3 float3 per ray
    DevBuf<int> bwd_hist, bwd_offs;             // backward: radix histograms
    DevBuf<unsigned> order_keys[2], order_vals[2];   // coherent ray order: radix sort ping-pong
    DevBuf<int> order_hist, order_offs;         // coherent ray order: radix histograms
    DevBuf<int> trace_work;                     // tracers: rays handed out so far
    DevBuf<int> mesh_off;                       // forest: triangle offsets per mesh (M+1)
    DevBuf<float4> tri_rec;                     // leaf records (v0 | id, e1, e2) in triangle order
    DevBuf<Lvl> lvl;                            // level state (one record)
    HostPinned host;
    LvlMirror mirror;                           // the level state as each eager level's leaf pass publishes it
};

}  // namespace pmkd

struct pm_ctx {
    int device = 0;
    std::string err;
    // geometry (borrowed)
    const float4 *p0 = nullptr, *p1 = nullptr, *p2 = nullptr;
    int ntris = 0;
    // tree
    bool has_tree = false;
    bool has_boxes = false;
    int num_nodes = 0, num_prim_indices = 0, num_leaf_ids = 0;
    int upper_levels = 0, lower_levels = 0;
    unsigned long long ref_reads = 0, ref_writes = 0, lower_prims = 0;   // work counts of the last build
    pmkd::NodeArrays nodes;
    pmkd::DevBuf<int> prim_indices;
    pmkd::DevBuf<int> leaf_ids;
    pmkd::DevBuf<int4> packed;      // trace view: packed nodes
    pmkd::DevBuf<float4> leaf_tris; // trace view: leaf-ordered triangles
    pmkd::DevBuf<float4> woop;      // Woop matrices (per leaf slot or per triangle), per call
    pmkd::BuildScratch scratch;
    std::vector<long long> level_refs;
    int num_roots = 1;                 // meshes in the last build (roots 0..num_roots-1)
    std::vector<int> mesh_offsets;     // triangle offsets per mesh (host copy)
    pmkd::BuildSizes last_sizes;       // what the last eager build reached
    long long last_ntris = -1;         // its triangle count (-1: no build yet)
    bool planned_build = false;        // the last build ran with capacities from the one before it
    // captured batch (build + trace replayed as one graph launch)
    hipGraphExec_t graph_exec = nullptr;
    hipStream_t capture_stream = nullptr;
    hipStream_t graph_last_stream = nullptr;
    pmkd::BuildSizes graph_caps;
    std::vector<std::pair<void **, void *>> graph_ptrs;   // buffers baked into the graph
    pmkd::DevBuf<int> graph_off;       // the captured batch's mesh offsets (its own copy)
    const float4 *graph_p0 = nullptr, *graph_p1 = nullptr, *graph_p2 = nullptr;
    int graph_n = 0, graph_M = 0;
    bool graph_boxes = false;
    std::vector<int> graph_offsets;
    bool graph_pending = false;        // a replay ran; host-side tree sizes not read back yet
    int num_cus = 0;                   // compute units of the device (0 = not queried yet)
    const int *trace_fail = nullptr;   // while recording: the build status the trace checks
    ~pm_ctx();
};

namespace pmkd {
// pm_build.hip
void build_tree(pm_ctx *ctx, const float4 *p0, const float4 *p1, const float4 *p2, int n,
                const int *mesh_offsets, int M, const pm_build_params &prm, hipStream_t s);
// record build + trace of a batch into a graph owned by ctx (see pm_batch_graph_capture)
void capture_batch(pm_ctx *ctx, const float4 *p0, const float4 *p1, const float4 *p2, int n,
                   const int *mesh_offsets, int M, const float4 *org, const float4 *dir, const int *ray_mesh,
                   int nrays, pm_hit *out, int algo, const pm_build_params &bp, const pm_trace_params &tp,
                   BuildSizes &caps, hipStream_t s);
// replay the captured batch on stream s (checks its buffers were not regrown)
void launch_batch(pm_ctx *ctx, hipStream_t s);
// after a replay: read the level state back and update ctx's host-side sizes;
// returns the build status (0 or a PM_ERR_*)
int sync_graph_state(pm_ctx *ctx, hipStream_t s);
void destroy_graph(pm_ctx *ctx);
// pm_trace.hip
void build_trace_view(pm_ctx *ctx, hipStream_t s);
// the two halves of the trace view: packed nodes, and the leaf slots'
// triangles (plus their areas when `areas` is given; `check` verifies the
// prim indices, needed for uploaded trees only).  With n_dev / m_dev the
// count is read on the device and ctx's count is its upper bound.  `rec`:
// the triangles' leaf records in triangle order (3 float4 each), copied
// instead of formed from the vertex arrays.
void build_node_view(pm_ctx *ctx, hipStream_t s, const int *n_dev = nullptr);
void build_leaf_view(pm_ctx *ctx, float *areas, bool check, hipStream_t s, const int *m_dev = nullptr,
                     const float4 *rec = nullptr);
void trace_canvas(pm_ctx *ctx, int algo, const float4 *org, const float4 *dir, int nrays, const float4 *n0,
                  const float4 *n1, const float4 *n2, const uchar4 *c0, const pm_light &light, int max_bounces,
                  uchar4 *canvas, const pm_trace_params &prm, hipStream_t s);
// ext (optional): the work counter and sort buffers to use instead of the
// context's (nothing is allocated then)
struct StaticScratch;
void trace_rays(pm_ctx *ctx, int algo, const float4 *org, const float4 *dir, int nrays,
                const int *ray_mesh, pm_hit *out, const pm_trace_params &prm, int4 *counters,
                hipStream_t s, const StaticScratch *ext = nullptr);
// pm_photon.hip: returns the number of continuation rays
int photon_bounce(pm_ctx *ctx, const float4 *org, const float4 *dir, const pm_hit *hits, int n,
                  int absorb_from, unsigned long long seed, float offset, float4 *out_org,
                  float4 *out_dir, int *out_src, int *out_count, hipStream_t s);
void photon_emit(pm_ctx *ctx, const float4 *l0, const float4 *l1, const float4 *l2, int m, int n,
                 unsigned long long seed, float offset, float4 *out_org, float4 *out_dir, hipStream_t s);
// pm_camera.hip: inv_view_proj (16) and cam_pos (3) are host arrays
void create_rays(const float *inv_view_proj, const float *cam_pos, int width, int height,
                 unsigned long long seed, int jitter, float4 *org, float4 *dir, hipStream_t s);
// pm_backward.hip
void trace_backward(pm_ctx *ctx, const float4 *p0, const float4 *p1, const float4 *p2, int ntris,
                    const float4 *org, const float4 *dir, const pm_hit *hits, int n,
                    const float *grad_t, const float *grad_u, const float *grad_v, const int *faces,
                    int num_vertices, float *grad_vertices, float *grad_org, float *grad_dir,
                    hipStream_t s);
// pm_geometry.hip
void autotune(pm_ctx *ctx, const float4 *p0, const float4 *p1, const float4 *p2, int n,
              const int *mesh_offsets, int M, const float4 *org, const float4 *dir,
              const int *ray_mesh, int nrays, const pm_build_params &base, float builds_per_trace,
              int reps, pm_tuning &out, hipStream_t s);
void osc_surface(float4 *vert, float4 *norm, int d, float height, float time, float amp, float freq,
                 hipStream_t s);
void radial_wave(const float4 *unit, int nv, float frame, const float *wave, float4 *out, hipStream_t s);
void add_mesh_attributes(pm_ctx *ctx, const float4 *normals, const uchar4 *colors, int nv, const unsigned *idx,
                         int nidx, const float *model, float4 *n0, float4 *n1, float4 *n2, uchar4 *c0,
                         uchar4 *c1, uchar4 *c2, hipStream_t s);
void add_mesh(pm_ctx *ctx, const float4 *verts, int nv, const unsigned *idx, int nidx,
              const float *model, float4 *p0, float4 *p1, float4 *p2, hipStream_t s);
// device-wide exclusive scan of n ints; out has n+1 entries (out[n] = total)
// exclusive scans with out[n] = total; with n_dev the length is
// n_mult * *n_dev (read on the device) and n is its upper bound
void scan_exclusive(const int *in, int *out, int n, BuildScratch &sc, hipStream_t s,
                    const int *n_dev = nullptr, int n_mult = 1);
void scan_exclusive2(const int *in0, int *out0, const int *in1, int *out1, int n, BuildScratch &sc,
                     hipStream_t s, const int *n_dev = nullptr, int n_mult = 1);
// block-sum buffers of every scan level for scans of up to n elements
void reserve_scan_tmp(BuildScratch &sc, long long n, hipStream_t s);
// pm_order.hip: stable LSD radix sort of n (key, value) pairs on key bits
// [0, bits), 8 bits per pass; keys[0]/vals[0] hold the input (vals unread
// with iota: the values are then the input indices; hist already holds the
// first digit's tile histograms with first_hist_ready); hist/offs need
// radix_hist_size(n) ints.  Returns which buffer holds the result.
int radix_sort_pairs(unsigned *const keys[2], unsigned *const vals[2], int n, int bits, int *hist, int *offs,
                     bool first_hist_ready, bool iota, BuildScratch &sc, hipStream_t s);
size_t radix_hist_size(long long n);
// coherent ray order: whether a trace of nrays rays uses it, its buffers
// (reserved ahead of a graph capture), and the order itself (device array:
// the ray indices in the order to trace them, valid until the next call on
// ctx)
bool ray_order_applies(const pm_trace_params &prm, int nrays);
void reserve_ray_order(pm_ctx *ctx, int nrays, hipStream_t s);
// the sort's buffers when they come from the caller instead of the context:
// keys/vals (nrays + 1 each), hist/offs (radix_hist_size(nrays) each)
struct OrderScratch {
    unsigned *keys[2], *vals[2];
    int *hist, *offs;
};
const unsigned *ray_order(pm_ctx *ctx, const float4 *org, const float4 *dir, const int *ray_root, int nrays,
                          hipStream_t s, const OrderScratch *ext = nullptr);
// pm_static.hip: the allocation-free trace's scratch (a file-scope device
// array), looked up once per device; rays the coherent order can take there
constexpr int NOALLOC_MAX_ORDERED_RAYS = PM_NOALLOC_MAX_ORDERED_RAYS;
struct StaticScratch {
    int *work;
    OrderScratch order;
};
// the calling thread's current device's scratch (PM_ERR_HIP if the symbol
// lookup fails); call with that device current
const StaticScratch &static_scratch(int device);
}  // namespace pmkd
