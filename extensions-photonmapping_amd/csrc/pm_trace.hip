// pm_trace.hip -- kd-restart, short-stack and brute-force traversal for
// gfx950 (wave64).  One lane per ray; the short stack lives in LDS,
// interleaved by lane so a wave's 64 accesses hit 64 different banks.
//
// Semantics (first hit of each ray):
//   kd-restart   Raytracer.cu:57-106 (TraceNode), :165-262 (KDRestart)
//   short-stack  ShortStack.cu:58-94 (TraceNode), :96-194, ShortStack.h:43-98
//   brute force  BruteTracer.cu:54-96 (Exhaustive)
//
// The tracers read a trace view built once per tree (build_trace_view):
//   packed[node]   int4 {split bits | prim index, left | bitmap lo,
//                        right | bitmap hi, axis}: one 16-byte load per visit
//   leaf_tris[3k]  float4 v0 (w = triangle id bits), e1 = v1 - v0,
//                  e2 = v2 - v0 for leaf slot k: a leaf's triangles are
//                  contiguous and need no prim-index indirection.
// e1/e2 are the same IEEE subtractions Moller-Trumbore performs, so hits are
// bit-identical to testing the original vertices.
#include <algorithm>
#include <climits>

#include "pm_device.h"
#include "pm_internal.h"

namespace pmkd {

namespace {

constexpr int TB = 256;   // threads per block (4 waves)

__device__ __forceinline__ pm_hit miss_record() {
    pm_hit h;
    h.id = -1; h.t = __int_as_float(0x7f800000); h.u = 0.0f; h.v = 0.0f;
    return h;
}

// Moller-Trumbore (IRayTracer.h:111-145) from precomputed edges.
__device__ __forceinline__ bool mt_edges(f3 v0, f3 e1, f3 e2, f3 o, f3 d, float &th, float &tu,
                                         float &tv) {
    const f3 p = cross3(d, e2);
    const f3 t = sub3(o, v0);
    const f3 q = cross3(t, e1);
    const float inv = 1.0f / dot3(p, e1);
    const float hx = inv * dot3(q, e2);
    if (0.0f <= hx && hx <= th) {
        const float hy = inv * dot3(p, t);
        const float hz = inv * dot3(q, d);
        if (hy >= 0.0f && hz >= 0.0f && hy + hz <= 1.0f) {
            th = hx; tu = hy; tv = hz;
            return true;
        }
    }
    return false;
}

// Leaf test (Raytracer.cu:194-216): optional box test, then Moller-Trumbore
// over the leaf bitmap in ascending bit order; the next triangle's data is
// loaded before the current one is tested.
template <bool SKIP, bool COUNT>
__device__ __forceinline__ int test_leaf(const TreeDev &T, const int4 nd, int node, f3 o, f3 d,
                                         f3 inv, float &th, float &tu, float &tv, int4 &cnt) {
    if (COUNT) cnt.y++;
    if (SKIP) {
        if (COUNT) cnt.w++;
        if (!ray_box(o, inv, mk3(T.bmin[node]), mk3(T.bmax[node]))) return -1;
    }
    unsigned long long bm = ((unsigned long long)(unsigned)nd.z << 32) | (unsigned)nd.y;
    if (!bm) return -1;
    const float4 *L = T.leaf_tris + 3ll * nd.x;
    int i = __builtin_ctzll(bm);
    float4 a = L[3 * i], b = L[3 * i + 1], c = L[3 * i + 2];
    int hit = -1;
    for (;;) {
        bm &= bm - 1;
        const bool more = bm != 0;
        float4 na, nb, nc;
        if (more) {
            const int j = __builtin_ctzll(bm);
            na = L[3 * j]; nb = L[3 * j + 1]; nc = L[3 * j + 2];
        }
        if (COUNT) cnt.z++;
        if (mt_edges(mk3(a), mk3(b), mk3(c), o, d, th, tu, tv)) hit = __float_as_int(a.w);
        if (!more) break;
        a = na; b = nb; c = nc;
    }
    return hit;
}

template <bool SKIP, bool COUNT>
__global__ __launch_bounds__(TB) void kd_restart_kernel(TreeDev T, const float4 *__restrict__ org,
                                                        const float4 *__restrict__ dir,
                                                        const int *__restrict__ ray_root, int n,
                                                        pm_hit *__restrict__ out,
                                                        int4 *__restrict__ counters) {
    const int r = blockIdx.x * TB + threadIdx.x;
    if (r >= n) return;
    int4 cnt = make_int4(0, 0, 0, 0);
    pm_hit res = miss_record();
    const int root_id = ray_root ? ray_root[r] : 0;
    // no such mesh, or (captured batch) the build before this trace failed: a miss
    if (root_id < 0 || root_id >= T.num_roots || (T.fail && *T.fail)) {
        out[r] = res;
        if (COUNT) counters[r] = cnt;
        return;
    }
    const f3 o = mk3(org[r]);
    const f3 d = adjust_dir(mk3(dir[r]));
    const f3 inv = {1.0f / d.x, 1.0f / d.y, 1.0f / d.z};
    float th = 0.0f, tu = 0.0f, tv = 0.0f;
    const int4 root = T.packed[root_id];
    long long visits = 0;
    do {
        if (++visits > (long long)T.num_nodes + 1) break;
        float tNext = F_INF;
        int node = root_id;
        int4 nd = root;
        int steps = 0;
        bool bad = false;
        while (nd.w != LEAF) {
            if (++steps > MAX_DESCENT) { bad = true; break; }
            if (COUNT) cnt.x++;
            const float s = __int_as_float(nd.x);
            const float ori = comp3(o, nd.w), dd = comp3(inv, nd.w);
            const float tSplit = (s - ori) * dd;
            if (th < tSplit) {
                node = 0 < dd ? nd.y : nd.z;
                tNext = fminf(tSplit, tNext);
            } else {
                node = 0 < dd ? nd.z : nd.y;
            }
            nd = T.packed[node];
        }
        if (bad) break;
        th = tNext * 1.00001f;
        const int hit = test_leaf<SKIP, COUNT>(T, nd, node, o, d, inv, th, tu, tv, cnt);
        if (hit != -1) {
            res.id = hit; res.t = th; res.u = tu; res.v = tv;
            break;
        }
    } while (th < F_INF);
    out[r] = res;
    if (COUNT) counters[r] = cnt;
}

template <int SMAX, bool SKIP, bool COUNT>
__global__ __launch_bounds__(TB) void short_stack_kernel(TreeDev T, const float4 *__restrict__ org,
                                                         const float4 *__restrict__ dir,
                                                         const int *__restrict__ ray_root, int n,
                                                         int S, pm_hit *__restrict__ out,
                                                         int4 *__restrict__ counters) {
    __shared__ int stk_node[SMAX][TB];
    __shared__ float stk_t[SMAX][TB];
    const int r = blockIdx.x * TB + threadIdx.x;
    if (r >= n) return;
    const int lt = threadIdx.x;
    int4 cnt = make_int4(0, 0, 0, 0);
    pm_hit res = miss_record();
    const int root_id = ray_root ? ray_root[r] : 0;
    // no such mesh, or (captured batch) the build before this trace failed: a miss
    if (root_id < 0 || root_id >= T.num_roots || (T.fail && *T.fail)) {
        out[r] = res;
        if (COUNT) counters[r] = cnt;
        return;
    }
    const f3 o = mk3(org[r]);
    const f3 d = adjust_dir(mk3(dir[r]));
    const f3 inv = {1.0f / d.x, 1.0f / d.y, 1.0f / d.z};
    float th = 0.0f, tu = 0.0f, tv = 0.0f;
    int nxt = 0, count = 0;
    const int4 root = T.packed[root_id];
    long long visits = 0;
    do {
        if (++visits > (long long)T.num_nodes + 1) break;
        int node;
        float tNext;
        int4 nd;
        if (count == 0) {
            node = root_id; tNext = F_INF; nd = root;
        } else {                                     // Pop, ShortStack.h:68-72
            nxt = (nxt == 0 ? S : nxt) - 1;
            count--;
            node = stk_node[nxt][lt];
            tNext = stk_t[nxt][lt];
            nd = T.packed[node];
        }
        int steps = 0;
        bool bad = false;
        while (nd.w != LEAF) {
            if (++steps > MAX_DESCENT) { bad = true; break; }
            if (COUNT) cnt.x++;
            const float s = __int_as_float(nd.x);
            const float ori = comp3(o, nd.w), dd = comp3(inv, nd.w);
            const float tSplit = (s - ori) * dd;
            const int lower = 0 < dd ? nd.y : nd.z;
            const int upper = 0 < dd ? nd.z : nd.y;
            if (th < tSplit) {
                node = lower;
                if (tSplit < tNext) {                // Push, ShortStack.h:57-62
                    stk_node[nxt][lt] = upper;
                    stk_t[nxt][lt] = tNext;
                    nxt = nxt == S - 1 ? 0 : nxt + 1;
                    count = count == S ? S : count + 1;
                }
                tNext = fminf(tSplit, tNext);
            } else {
                node = upper;
            }
            nd = T.packed[node];
        }
        if (bad) break;
        th = tNext * 1.00001f;
        const int hit = test_leaf<SKIP, COUNT>(T, nd, node, o, d, inv, th, tu, tv, cnt);
        if (hit != -1) {
            res.id = hit; res.t = th; res.u = tu; res.v = tv;
            break;
        }
    } while (th < F_INF);
    out[r] = res;
    if (COUNT) counters[r] = cnt;
}

// Brute force: every lane tests every triangle in index order; triangles are
// staged through LDS one tile at a time (one tile = TB triangles, 12 KiB).
template <bool COUNT>
__global__ __launch_bounds__(TB) void brute_kernel(const float4 *__restrict__ P0,
                                                   const float4 *__restrict__ P1,
                                                   const float4 *__restrict__ P2, int ntris,
                                                   const float4 *__restrict__ org,
                                                   const float4 *__restrict__ dir, int n,
                                                   pm_hit *__restrict__ out,
                                                   int4 *__restrict__ counters) {
    __shared__ float4 s0[TB], s1[TB], s2[TB];
    const int r = blockIdx.x * TB + threadIdx.x;
    const bool live = r < n;
    f3 o = {0.0f, 0.0f, 0.0f}, d = {0.0f, 0.0f, 1.0f};
    if (live) { o = mk3(org[r]); d = mk3(dir[r]); }
    float th = F_INF, tu = 0.0f, tv = 0.0f;
    int hit = -1;
    for (int base = 0; base < ntris; base += TB) {
        const int k = base + threadIdx.x;
        __syncthreads();
        if (k < ntris) { s0[threadIdx.x] = P0[k]; s1[threadIdx.x] = P1[k]; s2[threadIdx.x] = P2[k]; }
        __syncthreads();
        const int m = min(TB, ntris - base);
        if (live) {
            for (int j = 0; j < m; ++j) {
                if (moller_trumbore(mk3(s0[j]), mk3(s1[j]), mk3(s2[j]), o, d, th, tu, tv)) hit = base + j;
            }
        }
    }
    if (!live) return;
    pm_hit res = miss_record();
    if (th < F_INF && hit != -1) { res.id = hit; res.t = th; res.u = tu; res.v = tv; }
    out[r] = res;
    if (COUNT) counters[r] = make_int4(0, 0, ntris, 0);
}

// ---- trace view
// n = n_dev ? *n_dev : n_host (a captured build passes its device node count)
__global__ __launch_bounds__(TB) void pack_nodes_kernel(const int8_t *__restrict__ info,
                                                        const float *__restrict__ split,
                                                        const int2 *__restrict__ children,
                                                        const int *__restrict__ prim_index,
                                                        const long long *__restrict__ bitmap, int n_host,
                                                        const int *n_dev, int4 *__restrict__ out) {
    const int n = n_dev ? *n_dev : n_host;
    for (int i = blockIdx.x * TB + threadIdx.x; i < n; i += gridDim.x * TB) {
        const int ax = info[i] & 3;
        if (ax) {
            const int2 c = children[i];
            out[i] = make_int4(__float_as_int(split[i]), c.x, c.y, ax);
        } else {
            const unsigned long long b = (unsigned long long)bitmap[i];
            out[i] = make_int4(prim_index[i], (int)(unsigned)(b & 0xffffffffull), (int)(unsigned)(b >> 32), LEAF);
        }
    }
}

// leaf slot k -> v0 (w = id bits), e1, e2; with `areas`, also the slot's
// triangle area 0.5 |e1 x e2| (CalcSurfaceArea, LowerTriangleMap.h:24-38,
// the same operations tri_area performs)
__device__ __forceinline__ void leaf_slot(const int *__restrict__ prim_indices, int k,
                                          const float4 *__restrict__ P0, const float4 *__restrict__ P1,
                                          const float4 *__restrict__ P2, int ntris, float4 *__restrict__ out,
                                          int *bad, float *__restrict__ areas) {
    const int t = prim_indices[k];
    if (t < 0 || t >= ntris) {
        if (bad) atomicOr(bad, 1);
        const float nan = __int_as_float(0x7fc00000);
        out[3ll * k] = make_float4(nan, nan, nan, __int_as_float(-1));
        out[3ll * k + 1] = make_float4(nan, nan, nan, 0.0f);
        out[3ll * k + 2] = make_float4(nan, nan, nan, 0.0f);
        return;
    }
    const f3 v0 = mk3(P0[t]), v1 = mk3(P1[t]), v2 = mk3(P2[t]);
    const f3 e1 = sub3(v1, v0), e2 = sub3(v2, v0);
    out[3ll * k] = make_float4(v0.x, v0.y, v0.z, __int_as_float(t));
    out[3ll * k + 1] = make_float4(e1.x, e1.y, e1.z, 0.0f);
    out[3ll * k + 2] = make_float4(e2.x, e2.y, e2.z, 0.0f);
    if (areas) {
        const f3 c = cross3(e1, e2);
        areas[k] = 0.5f * sqrtf(dot3(c, c));
    }
}

// m = m_dev ? *m_dev : m_host
__global__ __launch_bounds__(TB) void leaf_tris_kernel(const int *__restrict__ prim_indices, int m_host,
                                                       const int *m_dev, const float4 *__restrict__ P0,
                                                       const float4 *__restrict__ P1,
                                                       const float4 *__restrict__ P2, int ntris,
                                                       float4 *__restrict__ out, int *bad,
                                                       float *__restrict__ areas) {
    const int m = m_dev ? *m_dev : m_host;
    for (int k = blockIdx.x * TB + threadIdx.x; k < m; k += gridDim.x * TB)
        leaf_slot(prim_indices, k, P0, P1, P2, ntris, out, bad, areas);
}

__global__ void clear_int_kernel(int *p) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = 0;
}

template <int SMAX, bool SKIP, bool COUNT>
void launch_ss(const TreeDev &T, const float4 *org, const float4 *dir, const int *root, int n, int S,
               pm_hit *out, int4 *cnt, hipStream_t s) {
    const int blocks = (n + TB - 1) / TB;
    short_stack_kernel<SMAX, SKIP, COUNT><<<blocks, TB, 0, s>>>(T, org, dir, root, n, S, out, cnt);
}

template <bool SKIP, bool COUNT>
void dispatch(const TreeDev &T, int algo, const float4 *org, const float4 *dir, const int *root, int n,
              int S, pm_hit *out, int4 *cnt, hipStream_t s) {
    const int blocks = (n + TB - 1) / TB;
    if (algo == PM_TRACE_KD_RESTART) {
        kd_restart_kernel<SKIP, COUNT><<<blocks, TB, 0, s>>>(T, org, dir, root, n, out, cnt);
    } else {
        if (S <= 4) launch_ss<4, SKIP, COUNT>(T, org, dir, root, n, S, out, cnt, s);
        else if (S <= 8) launch_ss<8, SKIP, COUNT>(T, org, dir, root, n, S, out, cnt, s);
        else launch_ss<16, SKIP, COUNT>(T, org, dir, root, n, S, out, cnt, s);
    }
}

}  // namespace

void build_node_view(pm_ctx *ctx, hipStream_t s, const int *n_dev) {
    const int n = ctx->num_nodes;
    ctx->packed.reserve((size_t)n + 1, s);
    if (n > 0) {
        const int blocks = std::min((n + TB - 1) / TB, n_dev ? 4096 : INT_MAX);
        pack_nodes_kernel<<<blocks, TB, 0, s>>>(ctx->nodes.info.ptr, ctx->nodes.split.ptr, ctx->nodes.children.ptr,
                                                ctx->nodes.prim_index.ptr, ctx->nodes.bitmap.ptr, n, n_dev,
                                                ctx->packed.ptr);
        PM_LAUNCH_CHECK();
    }
}

void build_leaf_view(pm_ctx *ctx, float *areas, bool check, hipStream_t s, const int *m_dev) {
    const int m = ctx->num_prim_indices;
    ctx->leaf_tris.reserve(3 * (size_t)m + 3, s);
    if (m == 0) return;
    BuildScratch &sc = ctx->scratch;
    int *bad = nullptr;
    if (check) {
        sc.totals.reserve(8, s);
        sc.host.ensure(8);
        clear_int_kernel<<<1, 64, 0, s>>>(sc.totals.ptr);
        PM_LAUNCH_CHECK();
        bad = sc.totals.ptr;
    }
    const int blocks = std::min((m + TB - 1) / TB, m_dev ? 8192 : INT_MAX);
    leaf_tris_kernel<<<blocks, TB, 0, s>>>(ctx->prim_indices.ptr, m, m_dev, ctx->p0, ctx->p1, ctx->p2,
                                           ctx->ntris, ctx->leaf_tris.ptr, bad, areas);
    PM_LAUNCH_CHECK();
    if (check) {
        PM_HIP(hipMemcpyAsync(sc.host.ptr, sc.totals.ptr, sizeof(int), hipMemcpyDeviceToHost, s));
        PM_HIP(hipStreamSynchronize(s));
        if (sc.host.ptr[0]) throw Error{PM_ERR_ARG, "prim index outside the triangle array"};
    }
}

void build_trace_view(pm_ctx *ctx, hipStream_t s) {
    build_node_view(ctx, s);
    build_leaf_view(ctx, nullptr, true, s);
}

void trace_rays(pm_ctx *ctx, int algo, const float4 *org, const float4 *dir, int nrays,
                const int *ray_mesh, pm_hit *out, const pm_trace_params &prm, int4 *counters,
                hipStream_t s) {
    if (nrays == 0) return;
    if (algo == PM_TRACE_BRUTE) {
        const int tris = prm.brute_cap > 0 ? std::min(ctx->ntris, prm.brute_cap) : ctx->ntris;
        const int blocks = (nrays + TB - 1) / TB;
        if (counters)
            brute_kernel<true><<<blocks, TB, 0, s>>>(ctx->p0, ctx->p1, ctx->p2, tris, org, dir, nrays, out, counters);
        else
            brute_kernel<false><<<blocks, TB, 0, s>>>(ctx->p0, ctx->p1, ctx->p2, tris, org, dir, nrays, out, counters);
        PM_LAUNCH_CHECK();
        return;
    }
    TreeDev T;
    T.info = ctx->nodes.info.ptr;
    T.split = ctx->nodes.split.ptr;
    T.children = ctx->nodes.children.ptr;
    T.prim_index = ctx->nodes.prim_index.ptr;
    T.bitmap = ctx->nodes.bitmap.ptr;
    T.bmin = ctx->nodes.bmin.ptr;
    T.bmax = ctx->nodes.bmax.ptr;
    T.prim_indices = ctx->prim_indices.ptr;
    T.packed = ctx->packed.ptr;
    T.leaf_tris = ctx->leaf_tris.ptr;
    T.num_nodes = ctx->num_nodes;
    T.num_roots = ctx->num_roots;
    T.fail = ctx->trace_fail;
    const bool skip = prm.leaf_skipping && ctx->has_boxes;
    const int S = prm.stack_size;
    if (skip) {
        if (counters) dispatch<true, true>(T, algo, org, dir, ray_mesh, nrays, S, out, counters, s);
        else dispatch<true, false>(T, algo, org, dir, ray_mesh, nrays, S, out, counters, s);
    } else {
        if (counters) dispatch<false, true>(T, algo, org, dir, ray_mesh, nrays, S, out, counters, s);
        else dispatch<false, false>(T, algo, org, dir, ray_mesh, nrays, S, out, counters, s);
    }
    PM_LAUNCH_CHECK();
}

}  // namespace pmkd
