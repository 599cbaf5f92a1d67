// pm_photon.hip -- diffuse continuation rays for photon paths (the caller
// side of the trace: emission -> trace -> bounce -> trace ...).
//
// The reference has no photon tracer despite its name; its only secondary
// rays are the reflection continuations of the shading loop
// (Raytracer.cu:219-256, ColorKernels.h:43-78).  This driver makes one
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
constexpr unsigned long long PCG_MULT = 6364136223846793005ull;

struct Pcg {
    unsigned long long state, inc;
    __device__ Pcg(unsigned long long seed, unsigned long long stream) {
        inc = (stream << 1) | 1ull;
        state = 0ull;
        state = state * PCG_MULT + inc;
        state += seed;
        state = state * PCG_MULT + inc;
    }
    __device__ unsigned int next() {
        const unsigned long long old = state;
        state = old * PCG_MULT + inc;
        const unsigned int xs = (unsigned int)(((old >> 18) ^ old) >> 27);
        const unsigned int rot = (unsigned int)(old >> 59);
        return (xs >> rot) | (xs << ((32u - rot) & 31u));
    }
    // 24 random bits in [0, 1)
    __device__ float uniform() { return (float)(next() >> 8) * (1.0f / 16777216.0f); }
};

__global__ __launch_bounds__(TB) void alive_kernel(const pm_hit *__restrict__ hits, int n,
                                                   int absorb_from, int *__restrict__ flags) {
    const int i = blockIdx.x * TB + threadIdx.x;
    if (i >= n) return;
    const int id = hits[i].id;
    flags[i] = id >= 0 && id < absorb_from;
}

__device__ __forceinline__ f3 normalize3(f3 v) {
    const float inv = 1.0f / sqrtf(dot3(v, v));
    return {v.x * inv, v.y * inv, v.z * inv};
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
    const float r = sqrtf(u1);
    const float phi = 6.283185307179586f * u2;
    float sp, cp;
    sincosf(phi, &sp, &cp);
    const float x = r * cp, y = r * sp, z = sqrtf(fmaxf(0.0f, 1.0f - u1));
    const f3 a = fabsf(nrm.x) > 0.9f ? f3{0.0f, 1.0f, 0.0f} : f3{1.0f, 0.0f, 0.0f};
    const f3 t = normalize3(cross3(a, nrm));
    const f3 b = cross3(nrm, t);
    const f3 nd = normalize3({x * t.x + y * b.x + z * nrm.x, x * t.y + y * b.y + z * nrm.y,
                              x * t.z + y * b.z + z * nrm.z});
    out_org[j] = make_float4(pos.x, pos.y, pos.z, 1.0f);
    out_dir[j] = make_float4(nd.x, nd.y, nd.z, 0.0f);
    if (out_src) out_src[j] = i;
}

__global__ void copy_count_kernel(const int *__restrict__ offs, int n, int *__restrict__ out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *out = offs[n];
}

}  // namespace

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
