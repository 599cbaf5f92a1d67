// pm_photon.hip -- photon emission and diffuse continuation rays for photon
// paths (the caller side of the trace: emission -> trace -> bounce -> trace
// ...).
//
// The reference has no photon tracer despite its name; its only secondary
// rays are the reflection continuations of the shading loop
// (Raytracer.cu:219-256, ColorKernels.h:43-78).  Emission samples the light
// triangles by area (emit_kernel below).  The bounce makes one
// cosine-weighted diffuse continuation per surviving hit:
//   origin    = o + t d + offset * n   (n: geometric normal facing the ray)
//   direction = cosine-weighted about n (Malley: uniform disk, lifted)
// with PCG32 (O'Neill 2014) stream (seed, purpose 6, source ray index), and
// a stable compaction (exclusive scan of survivor flags) so the output order
// is the input order.
#include "pm_device.h"
#include "pm_internal.h"

namespace pmkd {

namespace {

constexpr int TB = 256;
__global__ __launch_bounds__(TB) void alive_kernel(const pm_hit *__restrict__ hits, int n,
                                                   int absorb_from, int *__restrict__ flags) {
    const int i = blockIdx.x * TB + threadIdx.x;
    if (i >= n) return;
    const int id = hits[i].id;
    flags[i] = id >= 0 && id < absorb_from;
}

__global__ __launch_bounds__(TB) void bounce_kernel(
    const float4 *__restrict__ P0, const float4 *__restrict__ P1, const float4 *__restrict__ P2,
    const float4 *__restrict__ org, const float4 *__restrict__ dir, const pm_hit *__restrict__ hits,
    int n, const int *__restrict__ flags, const int *__restrict__ offs, unsigned long long seed,
    float offset, float4 *__restrict__ out_org, float4 *__restrict__ out_dir,
    int *__restrict__ out_src) {
    const int i = blockIdx.x * TB + threadIdx.x;
    if (i >= n || !flags[i]) return;
    const int j = offs[i];
    const pm_hit h = hits[i];
    const f3 v0 = mk3(P0[h.id]), v1 = mk3(P1[h.id]), v2 = mk3(P2[h.id]);
    f3 nrm = normalize3(cross3(sub3(v1, v0), sub3(v2, v0)));
    const f3 o = mk3(org[i]), d = mk3(dir[i]);
    if (dot3(nrm, d) > 0.0f) nrm = {-nrm.x, -nrm.y, -nrm.z};
    const f3 pos = {o.x + h.t * d.x + offset * nrm.x, o.y + h.t * d.y + offset * nrm.y,
                    o.z + h.t * d.z + offset * nrm.z};
    Pcg rng(seed, (6ull << 40) | (unsigned long long)i);
    const float u1 = rng.uniform(), u2 = rng.uniform();
    const f3 nd = cosine_dir(nrm, u1, u2);
    out_org[j] = make_float4(pos.x, pos.y, pos.z, 1.0f);
    out_dir[j] = make_float4(nd.x, nd.y, nd.z, 0.0f);
    if (out_src) out_src[j] = i;
}

__global__ void copy_count_kernel(const int *__restrict__ offs, int n, int *__restrict__ out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *out = offs[n];
}

// Emission: the running FP32 sum of the light triangles' areas
// (tri_area, ascending), one thread: cdf[k] = area_0 + ... + area_k.
__global__ void light_cdf_kernel(const float4 *__restrict__ L0, const float4 *__restrict__ L1,
                                 const float4 *__restrict__ L2, int m, float *__restrict__ cdf) {
    if (threadIdx.x | blockIdx.x) return;
    float acc = 0.0f;
    for (int k = 0; k < m; ++k) {
        acc += tri_area(mk3(L0[k]), mk3(L1[k]), mk3(L2[k]));
        cdf[k] = acc;
    }
}

// One photon per thread, PCG32 stream (seed, purpose 5, photon index):
//   u0 picks light k, the first with u0 * total < cdf[k] (the last if none);
//   (u1, u2) a uniform point of that triangle,
//     p = (1 - s) v0 + s (1 - u2) v1 + s u2 v2 with s = sqrt(u1);
//   (u3, u4) a cosine-weighted direction about n = normalize(e1 x e2);
// the origin is p + offset * n.
__global__ __launch_bounds__(TB) void emit_kernel(const float4 *__restrict__ L0, const float4 *__restrict__ L1,
                                                  const float4 *__restrict__ L2, int m,
                                                  const float *__restrict__ cdf, int n, unsigned long long seed,
                                                  float offset, float4 *__restrict__ out_org,
                                                  float4 *__restrict__ out_dir) {
    const int i = blockIdx.x * TB + threadIdx.x;
    if (i >= n) return;
    Pcg rng(seed, (5ull << 40) | (unsigned long long)i);
    const float u0 = rng.uniform(), u1 = rng.uniform(), u2 = rng.uniform();
    const float u3 = rng.uniform(), u4 = rng.uniform();
    const float target = u0 * cdf[m - 1];
    int lo = 0, hi = m - 1;            // first k with target < cdf[k]; m - 1 if none
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (target < cdf[mid]) hi = mid; else lo = mid + 1;
    }
    const f3 v0 = mk3(L0[lo]), v1 = mk3(L1[lo]), v2 = mk3(L2[lo]);
    const float s = sqrtf(u1);
    const float a = 1.0f - s, b = s * (1.0f - u2), c = s * u2;
    const f3 nrm = normalize3(cross3(sub3(v1, v0), sub3(v2, v0)));
    const f3 p = {a * v0.x + b * v1.x + c * v2.x, a * v0.y + b * v1.y + c * v2.y, a * v0.z + b * v1.z + c * v2.z};
    const f3 d = cosine_dir(nrm, u3, u4);
    out_org[i] = make_float4(p.x + offset * nrm.x, p.y + offset * nrm.y, p.z + offset * nrm.z, 1.0f);
    out_dir[i] = make_float4(d.x, d.y, d.z, 0.0f);
}

}  // namespace

void photon_emit(pm_ctx *ctx, const float4 *l0, const float4 *l1, const float4 *l2, int m, int n,
                 unsigned long long seed, float offset, float4 *out_org, float4 *out_dir, hipStream_t s) {
    if (n == 0) return;
    BuildScratch &sc = ctx->scratch;
    sc.light_cdf.reserve((size_t)m, s);
    light_cdf_kernel<<<1, 64, 0, s>>>(l0, l1, l2, m, sc.light_cdf.ptr);
    PM_LAUNCH_CHECK();
    emit_kernel<<<(n + TB - 1) / TB, TB, 0, s>>>(l0, l1, l2, m, sc.light_cdf.ptr, n, seed, offset, out_org,
                                                 out_dir);
    PM_LAUNCH_CHECK();
}

int photon_bounce(pm_ctx *ctx, const float4 *org, const float4 *dir, const pm_hit *hits, int n,
                  int absorb_from, unsigned long long seed, float offset, float4 *out_org,
                  float4 *out_dir, int *out_src, int *out_count, hipStream_t s) {
    BuildScratch &sc = ctx->scratch;
    sc.bounce_flags.reserve((size_t)n + 1, s);
    sc.bounce_offs.reserve((size_t)n + 1, s);
    if (n > 0) {
        alive_kernel<<<(n + TB - 1) / TB, TB, 0, s>>>(hits, n, absorb_from, sc.bounce_flags.ptr);
        PM_LAUNCH_CHECK();
    }
    scan_exclusive(sc.bounce_flags.ptr, sc.bounce_offs.ptr, n, sc, s);
    if (n > 0) {
        bounce_kernel<<<(n + TB - 1) / TB, TB, 0, s>>>(ctx->p0, ctx->p1, ctx->p2, org, dir, hits, n,
                                                       sc.bounce_flags.ptr, sc.bounce_offs.ptr, seed,
                                                       offset, out_org, out_dir, out_src);
        PM_LAUNCH_CHECK();
    }
    if (out_count) {
        copy_count_kernel<<<1, 64, 0, s>>>(sc.bounce_offs.ptr, n, out_count);
        PM_LAUNCH_CHECK();
    }
    sc.host.ensure(8);
    PM_HIP(hipMemcpyAsync(sc.host.ptr, sc.bounce_offs.ptr + n, sizeof(int), hipMemcpyDeviceToHost, s));
    PM_HIP(hipStreamSynchronize(s));
    return sc.host.ptr[0];
}

}  // namespace pmkd
