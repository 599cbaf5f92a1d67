// pm_device.h -- device-side helpers shared by the build and trace kernels.
//
// All float arithmetic is written in the reference's operation order and the
// library is compiled with -ffp-contract=off and correctly rounded fp32
// division/sqrt, so every value is the IEEE binary32 result of the same
// sequence of operations the reference's host branch performs.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pmkd {

// Utils.h:18: fInfinity is the integer 0x7f800000 used as a float value.
constexpr float F_INF = 2139095040.0f;
constexpr int LEAF = 0, AX_X = 1, AX_Y = 2, AX_Z = 3;  // KDNode.h:62-65
constexpr int MAX_DESCENT = 4096;

struct f3 { float x, y, z; };

__device__ __forceinline__ f3 mk3(const float4 &a) { return {a.x, a.y, a.z}; }
__device__ __forceinline__ f3 sub3(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 cross3(f3 a, f3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float comp3(f3 a, int axis) {
    return axis == AX_X ? a.x : (axis == AX_Y ? a.y : a.z);
}
__device__ __forceinline__ float comp4(const float4 &a, int axis) {
    return axis == AX_X ? a.x : (axis == AX_Y ? a.y : a.z);
}

// IEEE total order key: signed integer compare of keys orders floats (-0 < +0).
__device__ __forceinline__ int fkey(float f) {
    int b = __float_as_int(f);
    return b >= 0 ? b : (b ^ 0x7fffffff);
}
__device__ __forceinline__ float unkey(int k) {
    return __int_as_float(k >= 0 ? k : (k ^ 0x7fffffff));
}
__device__ __forceinline__ float omin(float a, float b) { return fkey(b) < fkey(a) ? b : a; }
__device__ __forceinline__ float omax(float a, float b) { return fkey(b) > fkey(a) ? b : a; }

// IRayTracer.h:60-65
__device__ __forceinline__ f3 adjust_dir(f3 d) {
    const float e = 0x1p-80f;
    d.x = fabsf(d.x) > e ? d.x : copysignf(e, d.x);
    d.y = fabsf(d.y) > e ? d.y : copysignf(e, d.y);
    d.z = fabsf(d.z) > e ? d.z : copysignf(e, d.z);
    return d;
}

// The IEEE quotient 1/x (correctly rounded), cheaply: for |x| in
// [2^-126, 2^126) the hardware reciprocal followed by one fma correction,
// y = y0 + y0 (1 - x y0), is the correctly rounded 1/x on gfx950 -- checked
// against 1.0f / x for every float of that range (tests/cpp/rcp_exhaustive.hip);
// outside it the full division.
__device__ __forceinline__ float rcp_rn(float x) {
    const float ax = fabsf(x);
    if (ax >= 0x1p-126f && ax < 0x1p126f) {
        const float y0 = __builtin_amdgcn_rcpf(x);
        const float e = __builtin_fmaf(-x, y0, 1.0f);
        return __builtin_fmaf(e, y0, y0);
    }
    return 1.0f / x;
}

// Moller-Trumbore as IRayTracer.h:111-145 (IEEE 1/det).  Later prims win
// equal t because the acceptance test is t <= th.
__device__ __forceinline__ bool moller_trumbore(f3 v0, f3 v1, f3 v2, f3 o, f3 d, float &th,
                                                float &tu, float &tv) {
    const f3 e1 = sub3(v1, v0);
    const f3 e2 = sub3(v2, v0);
    const f3 p = cross3(d, e2);
    const f3 t = sub3(o, v0);
    const f3 q = cross3(t, e1);
    const float inv = rcp_rn(dot3(p, e1));
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

// IRayTracer.h:147-168, invDir = true
__device__ __forceinline__ bool ray_box(f3 o, f3 inv, f3 bmin, f3 bmax) {
    bmin = {(bmin.x - o.x) * inv.x, (bmin.y - o.y) * inv.y, (bmin.z - o.z) * inv.z};
    bmax = {(bmax.x - o.x) * inv.x, (bmax.y - o.y) * inv.y, (bmax.z - o.z) * inv.z};
    float nearA = fminf(bmin.x, bmax.x);
    nearA = fmaxf(nearA, fminf(bmin.y, bmax.y));
    nearA = fmaxf(nearA, fminf(bmin.z, bmax.z));
    float farA = fmaxf(bmax.x, bmin.x);
    farA = fminf(farA, fmaxf(bmin.y, bmax.y));
    farA = fminf(farA, fmaxf(bmin.z, bmax.z));
    return nearA <= farA;
}

// LowerTriangleMap.h:32-36
__device__ __forceinline__ float tri_area(f3 v0, f3 v1, f3 v2) {
    const f3 c = cross3(sub3(v1, v0), sub3(v2, v0));
    return 0.5f * sqrtf(dot3(c, c));
}

// Triangle/box step 3 of Akenine-Moller (IntersectionTests.h:20-160,378-395):
// the nine edge x axis separating axes, each interval widened by GAMMA and
// EPSILON.  The a0X/a1X/a2X groups are one 2-D test on the (y,z), (z,x) and
// (x,y) components.
__device__ __forceinline__ bool sat_edge(float au, float aw, float bu, float bw, float cu, float cw,
                                         float hu, float hw) {
    const float p0 = aw * bu - au * bw;
    const float p1 = (bu - au) * cw - (bw - aw) * cu;
    float r = hw * fabsf(bu - au) + hu * fabsf(bw - aw) + 0.00001f;
    r *= 1.0001f;
    return !((p0 > r && p1 > r) || (p0 < -r && p1 < -r));
}

__device__ __forceinline__ bool sat_plane(float u0, float w0, float u1, float w1, float u2, float w2,
                                          float hu, float hw) {
    bool ret = true;
    ret &= sat_edge(u0, w0, u1, w1, u2, w2, hu, hw);
    ret &= sat_edge(u1, w1, u2, w2, u0, w0, hu, hw);
    ret &= sat_edge(u2, w2, u0, w0, u1, w1, hu, hw);
    return ret;
}

__device__ __forceinline__ bool tri_box_step3(f3 v0, f3 v1, f3 v2, f3 bmin, f3 bmax) {
    const f3 h = {(bmax.x - bmin.x) * 0.5f, (bmax.y - bmin.y) * 0.5f, (bmax.z - bmin.z) * 0.5f};
    const f3 c = {bmin.x + h.x, bmin.y + h.y, bmin.z + h.z};
    v0 = sub3(v0, c); v1 = sub3(v1, c); v2 = sub3(v2, c);
    bool ret = true;
    ret &= sat_plane(v0.y, v0.z, v1.y, v1.z, v2.y, v2.z, h.y, h.z);
    ret &= sat_plane(v0.z, v0.x, v1.z, v1.x, v2.z, v2.x, h.z, h.x);
    ret &= sat_plane(v0.x, v0.y, v1.x, v1.y, v2.x, v2.y, h.x, h.y);
    return ret;
}

// WoopTransformationMatrix (IntersectionTests.h:437-467)
__device__ __forceinline__ void woop_matrix(f3 a, f3 b, f3 c, float4 &m0, float4 &m1, float4 &m2) {
    const f3 i0 = sub3(a, c), i1 = sub3(b, c), i2 = sub3(cross3(i0, i1), c);
    const float det = i0.x * (i1.y * i2.z - i1.z * i2.y) + i1.x * (i2.y * i0.z - i2.z * i0.y) +
                      i2.x * (i0.y * i1.z - i0.z * i1.y);
    const float inv = 1.0f / det;
    const f3 t0 = {(i1.y * i2.z - i1.z * i2.y) * inv, (i2.x * i1.z - i2.z * i1.x) * inv,
                   (i1.x * i2.y - i1.y * i2.x) * inv};
    const f3 t1 = {(i2.y * i0.z - i2.z * i0.y) * inv, (i0.x * i2.z - i0.z * i2.x) * inv,
                   (i2.x * i0.y - i2.y * i0.x) * inv};
    const f3 t2 = {(i0.y * i1.z - i0.z * i1.y) * inv, (i1.x * i0.z - i1.z * i0.x) * inv,
                   (i0.x * i1.y - i0.y * i1.x) * inv};
    m0 = make_float4(t0.x, t0.y, t0.z, dot3(t0, c));
    m1 = make_float4(t1.x, t1.y, t1.z, dot3(t1, c));
    m2 = make_float4(t2.x, t2.y, t2.z, dot3(t2, c));
}

// IRayTracer::Woop with WoopLambda / WoopUV (IRayTracer.h:78-109)
__device__ __forceinline__ bool woop_test(const float4 &m0, const float4 &m1, const float4 &m2, f3 o, f3 d,
                                          float &th, float &tu, float &tv) {
    const f3 n = mk3(m2);
    const float lambda = -(dot3(n, o) - m2.w) / dot3(n, d);
    if (0.0f <= lambda && lambda <= th) {
        const f3 a = mk3(m0), b = mk3(m1);
        const float u = lambda * dot3(a, d) + dot3(a, o) - m0.w;
        const float v = lambda * dot3(b, d) + dot3(b, o) - m1.w;
        if (u >= -0.0f && v >= -0.0f && u + v <= 1.0f) {
            th = lambda; tu = u; tv = v;
            return true;
        }
    }
    return false;
}

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// PCG32 (O'Neill 2014: pcg32_srandom_r / pcg32_random_r), one stream per
// (purpose, element): stream = purpose << 40 | element.
struct Pcg {
    unsigned long long state, inc;
    static constexpr unsigned long long MULT = 6364136223846793005ull;
    __device__ Pcg(unsigned long long seed, unsigned long long stream) {
        inc = (stream << 1) | 1ull;
        state = 0ull;
        state = state * MULT + inc;
        state += seed;
        state = state * MULT + inc;
    }
    __device__ unsigned int next() {
        const unsigned long long old = state;
        state = old * MULT + inc;
        const unsigned int xs = (unsigned int)(((old >> 18) ^ old) >> 27);
        const unsigned int rot = (unsigned int)(old >> 59);
        return (xs >> rot) | (xs << ((32u - rot) & 31u));
    }
    // 24 random bits in [0, 1)
    __device__ float uniform() { return (float)(next() >> 8) * (1.0f / 16777216.0f); }
};

// v * (1 / |v|), the host branch of cutil_math's normalize (rsqrtf = 1/sqrtf)
__device__ __forceinline__ f3 normalize3(f3 v) {
    const float inv = 1.0f / sqrtf(dot3(v, v));
    return {v.x * inv, v.y * inv, v.z * inv};
}

// Cosine-weighted direction about the unit normal n (Malley: a uniform point
// (r cos phi, r sin phi) of the unit disk, r = sqrt(u1), phi = 2 pi u2,
// lifted to the hemisphere) in the frame t = normalize(a x n), b = n x t.
__device__ __forceinline__ f3 cosine_dir(f3 n, float u1, float u2) {
    const float r = sqrtf(u1);
    const float phi = 6.283185307179586f * u2;
    float sp, cp;
    sincosf(phi, &sp, &cp);
    const float x = r * cp, y = r * sp, z = sqrtf(fmaxf(0.0f, 1.0f - u1));
    const f3 a = fabsf(n.x) > 0.9f ? f3{0.0f, 1.0f, 0.0f} : f3{1.0f, 0.0f, 0.0f};
    const f3 t = normalize3(cross3(a, n));
    const f3 b = cross3(n, t);
    return normalize3({x * t.x + y * b.x + z * n.x, x * t.y + y * b.y + z * n.y, x * t.z + y * b.z + z * n.z});
}

}  // namespace pmkd
