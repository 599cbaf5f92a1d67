// pm_trace.hip -- kd-restart, short-stack and brute-force traversal for
// gfx950 (wave64).  One lane per ray; the short stack lives in LDS,
// interleaved by lane so a wave's 64 accesses hit 64 different banks.
//
// Semantics (first hit of each ray):
//   kd-restart   Raytracer.cu:57-106 (TraceNode), :165-262 (KDRestart)
//   short-stack  ShortStack.cu:58-94 (TraceNode), :96-194, ShortStack.h:43-98
//   brute force  BruteTracer.cu:54-96 (Exhaustive)
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

// Leaf test (Raytracer.cu:194-216): optional box test, then Moller-Trumbore
// over the leaf bitmap in ascending bit order.
template <bool SKIP, bool COUNT>
__device__ __forceinline__ int test_leaf(const TreeDev &T, const float4 *__restrict__ P0,
                                         const float4 *__restrict__ P1,
                                         const float4 *__restrict__ P2, int node, f3 o, f3 d,
                                         f3 inv, float &th, float &tu, float &tv, int4 &cnt) {
    if (COUNT) cnt.y++;
    if (SKIP) {
        if (COUNT) cnt.w++;
        if (!ray_box(o, inv, mk3(T.bmin[node]), mk3(T.bmax[node]))) return -1;
    }
    const int pi = T.prim_index[node];
    unsigned long long bm = (unsigned long long)T.bitmap[node];
    int hit = -1;
    while (bm) {
        const int i = __builtin_ctzll(bm);
        const int prim = T.prim_indices[pi + i];
        if (COUNT) cnt.z++;
        if (moller_trumbore(mk3(P0[prim]), mk3(P1[prim]), mk3(P2[prim]), o, d, th, tu, tv)) hit = prim;
        bm &= bm - 1;
    }
    return hit;
}

template <bool SKIP, bool COUNT>
__global__ __launch_bounds__(TB) void kd_restart_kernel(TreeDev T, const float4 *__restrict__ P0,
                                                        const float4 *__restrict__ P1,
                                                        const float4 *__restrict__ P2,
                                                        const float4 *__restrict__ org,
                                                        const float4 *__restrict__ dir, int n,
                                                        pm_hit *__restrict__ out,
                                                        int4 *__restrict__ counters) {
    const int r = blockIdx.x * TB + threadIdx.x;
    if (r >= n) return;
    int4 cnt = make_int4(0, 0, 0, 0);
    const f3 o = mk3(org[r]);
    const f3 d = adjust_dir(mk3(dir[r]));
    const f3 inv = {1.0f / d.x, 1.0f / d.y, 1.0f / d.z};
    float th = 0.0f, tu = 0.0f, tv = 0.0f;
    pm_hit res = miss_record();
    long long visits = 0;
    do {
        if (++visits > (long long)T.num_nodes + 1) break;
        float tNext = F_INF;
        int node = 0;
        int axis = T.info[0];
        int steps = 0;
        bool bad = false;
        while ((axis & 3) != LEAF) {
            if (++steps > MAX_DESCENT) { bad = true; break; }
            if (COUNT) cnt.x++;
            const float s = T.split[node];
            const int2 ch = T.children[node];
            const int a = axis & 3;
            const float ori = comp3(o, a), dd = comp3(inv, a);
            const float tSplit = (s - ori) * dd;
            if (th < tSplit) {
                node = 0 < dd ? ch.x : ch.y;
                tNext = fminf(tSplit, tNext);
            } else {
                node = 0 < dd ? ch.y : ch.x;
            }
            axis = T.info[node];
        }
        if (bad) break;
        th = tNext * 1.00001f;
        const int hit = test_leaf<SKIP, COUNT>(T, P0, P1, P2, node, o, d, inv, th, tu, tv, cnt);
        if (hit != -1) {
            res.id = hit; res.t = th; res.u = tu; res.v = tv;
            break;
        }
    } while (th < F_INF);
    out[r] = res;
    if (COUNT) counters[r] = cnt;
}

template <int SMAX, bool SKIP, bool COUNT>
__global__ __launch_bounds__(TB) void short_stack_kernel(TreeDev T, const float4 *__restrict__ P0,
                                                         const float4 *__restrict__ P1,
                                                         const float4 *__restrict__ P2,
                                                         const float4 *__restrict__ org,
                                                         const float4 *__restrict__ dir, int n,
                                                         int S, pm_hit *__restrict__ out,
                                                         int4 *__restrict__ counters) {
    __shared__ int stk_node[SMAX][TB];
    __shared__ float stk_t[SMAX][TB];
    const int r = blockIdx.x * TB + threadIdx.x;
    if (r >= n) return;
    const int lt = threadIdx.x;
    int4 cnt = make_int4(0, 0, 0, 0);
    const f3 o = mk3(org[r]);
    const f3 d = adjust_dir(mk3(dir[r]));
    const f3 inv = {1.0f / d.x, 1.0f / d.y, 1.0f / d.z};
    float th = 0.0f, tu = 0.0f, tv = 0.0f;
    int nxt = 0, count = 0;
    pm_hit res = miss_record();
    long long visits = 0;
    do {
        if (++visits > (long long)T.num_nodes + 1) break;
        int node;
        float tNext;
        if (count == 0) {
            node = 0; tNext = F_INF;
        } else {                                     // Pop, ShortStack.h:68-72
            nxt = (nxt == 0 ? S : nxt) - 1;
            count--;
            node = stk_node[nxt][lt];
            tNext = stk_t[nxt][lt];
        }
        int info = T.info[node];
        int steps = 0;
        bool bad = false;
        while ((info & 3) != LEAF) {
            if (++steps > MAX_DESCENT) { bad = true; break; }
            if (COUNT) cnt.x++;
            const float s = T.split[node];
            const int2 ch = T.children[node];
            const int a = info & 3;
            const float ori = comp3(o, a), dd = comp3(inv, a);
            const float tSplit = (s - ori) * dd;
            const int lower = 0 < dd ? ch.x : ch.y;
            const int upper = 0 < dd ? ch.y : ch.x;
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
            info = T.info[node];
        }
        if (bad) break;
        th = tNext * 1.00001f;
        const int hit = test_leaf<SKIP, COUNT>(T, P0, P1, P2, node, o, d, inv, th, tu, tv, cnt);
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

template <int SMAX, bool SKIP, bool COUNT>
void launch_ss(const TreeDev &T, const float4 *P0, const float4 *P1, const float4 *P2,
               const float4 *org, const float4 *dir, int n, int S, pm_hit *out, int4 *cnt,
               hipStream_t s) {
    const int blocks = (n + TB - 1) / TB;
    short_stack_kernel<SMAX, SKIP, COUNT><<<blocks, TB, 0, s>>>(T, P0, P1, P2, org, dir, n, S, out, cnt);
}

template <bool SKIP, bool COUNT>
void dispatch(const TreeDev &T, int algo, const float4 *P0, const float4 *P1, const float4 *P2,
              const float4 *org, const float4 *dir, int n, int S, pm_hit *out, int4 *cnt,
              hipStream_t s) {
    const int blocks = (n + TB - 1) / TB;
    if (algo == PM_TRACE_KD_RESTART) {
        kd_restart_kernel<SKIP, COUNT><<<blocks, TB, 0, s>>>(T, P0, P1, P2, org, dir, n, out, cnt);
    } else {
        if (S <= 4) launch_ss<4, SKIP, COUNT>(T, P0, P1, P2, org, dir, n, S, out, cnt, s);
        else if (S <= 8) launch_ss<8, SKIP, COUNT>(T, P0, P1, P2, org, dir, n, S, out, cnt, s);
        else launch_ss<16, SKIP, COUNT>(T, P0, P1, P2, org, dir, n, S, out, cnt, s);
    }
}

}  // namespace

void trace_rays(pm_ctx *ctx, int algo, const float4 *org, const float4 *dir, int nrays,
                pm_hit *out, const pm_trace_params &prm, int4 *counters, hipStream_t s) {
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
    T.num_nodes = ctx->num_nodes;
    const bool skip = prm.leaf_skipping && ctx->has_boxes;
    const int S = prm.stack_size;
    if (skip) {
        if (counters) dispatch<true, true>(T, algo, ctx->p0, ctx->p1, ctx->p2, org, dir, nrays, S, out, counters, s);
        else dispatch<true, false>(T, algo, ctx->p0, ctx->p1, ctx->p2, org, dir, nrays, S, out, counters, s);
    } else {
        if (counters) dispatch<false, true>(T, algo, ctx->p0, ctx->p1, ctx->p2, org, dir, nrays, S, out, counters, s);
        else dispatch<false, false>(T, algo, ctx->p0, ctx->p1, ctx->p2, org, dir, nrays, S, out, counters, s);
    }
    PM_LAUNCH_CHECK();
}

}  // namespace pmkd
