// pm_backward.hip -- gradients of the ray/triangle intersection.
//
// Forward (IRayTracer.h:111-145): e1 = v1-v0, e2 = v2-v0, p = d x e2,
// T = o - v0, q = T x e1, det = p.e1, inv = 1/det, t = inv (q.e2),
// u = inv (p.T), v = inv (q.d).  For every hit ray the backward kernel
// recomputes these in FP32 and applies the reverse-mode adjoint
//   a = q.e2, b = p.T, c = q.d
//   ga = gt inv, gb = gu inv, gc = gv inv, ginv = gt a + gu b + gv c,
//   gdet = -ginv inv inv,
//   gq = ga e2 + gc d, ge2 = ga q + gp x d, gp = gb T + gdet e1,
//   gT = gb p + e1 x gq, ge1 = gdet p + gq x T, gd = gc q + e2 x gp,
//   grad o = gT, grad d = gd, grad v1 = ge1, grad v2 = ge2,
//   grad v0 = -gT - ge1 - ge2
// (no counterpart in the reference; checked against FP64 central
// differences and an FP64 adjoint).  A record whose id is not a triangle of
// the given geometry, or whose triangle's face indices are outside
// [0, num_vertices), contributes nothing and fails the call (checked in the
// kernel, read back once at the end).  Per-vertex gradients are reduced
// deterministically: the 3 corner entries of every ray are keyed by vertex
// id, put in key order by a stable LSD radix sort (8-bit digits, wave-ballot
// multisplit ranks), and each vertex sums its entries sequentially in
// (ray, corner) order, so the result is bit-identical from run to run.
#include "pm_device.h"
#include "pm_internal.h"

namespace pmkd {

namespace {

constexpr int TB = 256;

__device__ __forceinline__ f3 scale3(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ f3 add3(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }

// one ray: corner gradients (3 x float3) and origin/direction gradients
__global__ __launch_bounds__(TB) void mt_backward_kernel(
    const float4 *__restrict__ P0, const float4 *__restrict__ P1, const float4 *__restrict__ P2,
    int ntris, const float4 *__restrict__ org, const float4 *__restrict__ dir,
    const pm_hit *__restrict__ hits, int n, const float *__restrict__ grad_t,
    const float *__restrict__ grad_u, const float *__restrict__ grad_v, const int *__restrict__ faces,
    unsigned miss_key, unsigned *__restrict__ keys, float *__restrict__ corner,
    float *__restrict__ grad_org, float *__restrict__ grad_dir, int *__restrict__ bad) {
    const int r = blockIdx.x * TB + threadIdx.x;
    if (r >= n) return;
    int id = hits[r].id;
    // a record that names no triangle of this geometry, or a hit triangle
    // whose corners name no vertex, contributes nothing and fails the call
    int vk[3] = {0, 0, 0};
    if (id >= ntris || id < -1) {
        atomicOr(bad, 1);
        id = -1;
    } else if (id >= 0) {
        for (int k = 0; k < 3; ++k) vk[k] = faces ? faces[3 * id + k] : 3 * id + k;
        if ((unsigned)vk[0] >= miss_key || (unsigned)vk[1] >= miss_key || (unsigned)vk[2] >= miss_key) {
            atomicOr(bad, 2);
            id = -1;
        }
    }
    const float gt = grad_t ? grad_t[r] : 0.0f;
    const float gu = grad_u ? grad_u[r] : 0.0f;
    const float gv = grad_v ? grad_v[r] : 0.0f;
    f3 g_o = {0.0f, 0.0f, 0.0f}, g_d = g_o, g0 = g_o, g1 = g_o, g2 = g_o;
    if (id >= 0) {
        const f3 v0 = mk3(P0[id]), v1 = mk3(P1[id]), v2 = mk3(P2[id]);
        const f3 o = mk3(org[r]), d = mk3(dir[r]);
        const f3 e1 = sub3(v1, v0), e2 = sub3(v2, v0), T = sub3(o, v0);
        const f3 p = cross3(d, e2), q = cross3(T, e1);
        const float det = dot3(p, e1);
        const float inv = 1.0f / det;
        const float a = dot3(q, e2), b = dot3(p, T), c = dot3(q, d);
        const float ga = gt * inv, gb = gu * inv, gc = gv * inv;
        const float ginv = gt * a + gu * b + gv * c;
        const float gdet = -ginv * inv * inv;
        const f3 gq = add3(scale3(e2, ga), scale3(d, gc));
        const f3 gp = add3(scale3(T, gb), scale3(e1, gdet));
        const f3 gT = add3(scale3(p, gb), cross3(e1, gq));
        const f3 ge1 = add3(scale3(p, gdet), cross3(gq, T));
        const f3 gd = add3(scale3(q, gc), cross3(e2, gp));
        const f3 ge2 = add3(scale3(q, ga), cross3(gp, d));
        g_o = gT; g_d = gd; g1 = ge1; g2 = ge2;
        g0 = {-gT.x - ge1.x - ge2.x, -gT.y - ge1.y - ge2.y, -gT.z - ge1.z - ge2.z};
    }
    const f3 gc3[3] = {g0, g1, g2};
    for (int k = 0; k < 3; ++k) {
        const long long e = 3ll * r + k;
        keys[e] = id >= 0 ? (unsigned)vk[k] : miss_key;
        corner[3 * e] = gc3[k].x; corner[3 * e + 1] = gc3[k].y; corner[3 * e + 2] = gc3[k].z;
    }
    if (grad_org) { grad_org[3 * r] = g_o.x; grad_org[3 * r + 1] = g_o.y; grad_org[3 * r + 2] = g_o.z; }
    if (grad_dir) { grad_dir[3 * r] = g_d.x; grad_dir[3 * r + 1] = g_d.y; grad_dir[3 * r + 2] = g_d.z; }
}

// per vertex: sequential sum of its entries in sorted (= ray, corner) order
__global__ __launch_bounds__(TB) void segment_sum_kernel(const unsigned *__restrict__ keys,
                                                         const unsigned *__restrict__ vals, int n,
                                                         unsigned miss_key,
                                                         const float *__restrict__ corner,
                                                         float *__restrict__ out) {
    const int i = blockIdx.x * TB + threadIdx.x;
    if (i >= n) return;
    const unsigned k = keys[i];
    if (k == miss_key || (i > 0 && keys[i - 1] == k)) return;
    float sx = 0.0f, sy = 0.0f, sz = 0.0f;
    for (int j = i; j < n && keys[j] == k; ++j) {
        const unsigned e = vals[j];
        sx += corner[3ull * e]; sy += corner[3ull * e + 1]; sz += corner[3ull * e + 2];
    }
    out[3ull * k] = sx; out[3ull * k + 1] = sy; out[3ull * k + 2] = sz;
}

__global__ void zero_kernel(float *p, long long n) {
    const long long i = (long long)blockIdx.x * TB + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

__global__ void clear_flag_kernel(int *p) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = 0;
}

}  // namespace

void trace_backward(pm_ctx *ctx, const float4 *P0, const float4 *P1, const float4 *P2, int ntris,
                    const float4 *org, const float4 *dir, const pm_hit *hits, int n,
                    const float *grad_t, const float *grad_u, const float *grad_v, const int *faces,
                    int num_vertices, float *grad_vertices, float *grad_org, float *grad_dir,
                    hipStream_t s) {
    BuildScratch &sc = ctx->scratch;
    const long long total = 3ll * num_vertices;
    if (total > 0) {
        zero_kernel<<<(int)((total + TB - 1) / TB), TB, 0, s>>>(grad_vertices, total);
        PM_LAUNCH_CHECK();
    }
    if (n == 0) return;
    if (3ll * n > (long long)INT32_MAX / 2) throw Error{PM_ERR_ARG, "too many rays for one backward call"};
    const int m = 3 * n;
    if (!faces && 3ll * ntris > num_vertices)
        throw Error{PM_ERR_ARG, "without faces, num_vertices must be >= 3 * triangles"};
    sc.totals.reserve(8, s);
    sc.host.ensure(8);
    clear_flag_kernel<<<1, 64, 0, s>>>(sc.totals.ptr);
    PM_LAUNCH_CHECK();
    const unsigned miss_key = (unsigned)num_vertices;
    sc.bwd_keys[0].reserve(m, s); sc.bwd_keys[1].reserve(m, s);
    sc.bwd_vals[0].reserve(m, s); sc.bwd_vals[1].reserve(m, s);
    sc.bwd_corner.reserve(3 * (size_t)m, s);
    mt_backward_kernel<<<(n + TB - 1) / TB, TB, 0, s>>>(P0, P1, P2, ntris, org, dir, hits, n, grad_t, grad_u,
                                                       grad_v, faces, miss_key, sc.bwd_keys[0].ptr,
                                                       sc.bwd_corner.ptr, grad_org, grad_dir, sc.totals.ptr);
    PM_LAUNCH_CHECK();
    int bits = 1;
    while (bits < 32 && ((unsigned long long)miss_key >> bits)) ++bits;
    sc.bwd_hist.reserve(radix_hist_size(m), s);
    sc.bwd_offs.reserve(radix_hist_size(m), s);
    reserve_scan_tmp(sc, (long long)radix_hist_size(m), s);
    unsigned *const K[2] = {sc.bwd_keys[0].ptr, sc.bwd_keys[1].ptr};
    unsigned *const V[2] = {sc.bwd_vals[0].ptr, sc.bwd_vals[1].ptr};
    // values = entry indices (3 r + corner): the sorted order within a vertex is (ray, corner)
    const int cur = radix_sort_pairs(K, V, m, bits, sc.bwd_hist.ptr, sc.bwd_offs.ptr, false, true, sc, s);
    segment_sum_kernel<<<(m + TB - 1) / TB, TB, 0, s>>>(K[cur], V[cur], m, miss_key,
                                                        sc.bwd_corner.ptr, grad_vertices);
    PM_LAUNCH_CHECK();
    PM_HIP(hipMemcpyAsync(sc.host.ptr, sc.totals.ptr, sizeof(int), hipMemcpyDeviceToHost, s));
    PM_HIP(hipStreamSynchronize(s));
    if (sc.host.ptr[0] & 1) throw Error{PM_ERR_ARG, "hit id outside [-1, triangles)"};
    if (sc.host.ptr[0] & 2) throw Error{PM_ERR_ARG, "face index outside [0, num_vertices)"};
}

}  // namespace pmkd
