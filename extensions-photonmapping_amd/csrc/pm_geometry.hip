// pm_geometry.hip -- geometry staging: indexed mesh -> triangle SoA, and
// the per-frame vertex deformers that feed a rebuild.
//
// AddMeshKernel (float4 vertex variant, GeometryList.cu:197-232): triangle
// id = index triple 3*id..3*id+2, each vertex transformed by the model
// matrix row by row, p_i = dot(row_i, v) evaluated left to right
// (Math/CUDA/Matrix.h:44-49).  One thread per triangle (the reference
// launches one thread per index, GeometryList.cu:331); indices outside the
// vertex array are reported instead of read.
#include "pm_device.h"
#include "pm_internal.h"

namespace pmkd {

namespace {

constexpr int TB = 256;

struct Mat4 { float4 r[4]; };

__device__ __forceinline__ float dot4(const float4 &a, const float4 &b) {
    return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w;
}

__global__ __launch_bounds__(TB) void add_mesh_kernel(const float4 *__restrict__ verts, int nv,
                                                      const unsigned *__restrict__ idx, int ntris,
                                                      Mat4 m, float4 *__restrict__ p0,
                                                      float4 *__restrict__ p1,
                                                      float4 *__restrict__ p2, int *bad) {
    const int id = blockIdx.x * TB + threadIdx.x;
    if (id >= ntris) return;
    const unsigned i0 = idx[3 * id], i1 = idx[3 * id + 1], i2 = idx[3 * id + 2];
    if (i0 >= (unsigned)nv || i1 >= (unsigned)nv || i2 >= (unsigned)nv) {
        atomicOr(bad, 1);
        return;
    }
    const float4 v[3] = {verts[i0], verts[i1], verts[i2]};
    float4 *out[3] = {p0, p1, p2};
    for (int k = 0; k < 3; ++k)
        out[k][id] = make_float4(dot4(m.r[0], v[k]), dot4(m.r[1], v[k]), dot4(m.r[2], v[k]), dot4(m.r[3], v[k]));
}

// UpdateSurface (OscCUDAMeshNode.cu:72-93): vertex (x, y) of a d x d grid,
// stored at y * d + x, gets height + A cos(f x + time) + A sin(f y + time)
// as its y coordinate and the normal (-A f sin(f x + time),
// 1 - nx^2 - nz^2, A f cos(f y + time), 0).
__global__ __launch_bounds__(TB) void osc_surface_kernel(float4 *__restrict__ vert,
                                                         float4 *__restrict__ norm, int d,
                                                         float height, float time, float amp,
                                                         float freq) {
    const int id = blockIdx.x * TB + threadIdx.x;
    if (id >= d * d) return;
    const int x = id % d, y = id / d;
    const float ax = freq * (float)x + time, ay = freq * (float)y + time;
    float4 v = vert[id];
    v.y = height + amp * cosf(ax);
    v.y += amp * sinf(ay);
    vert[id] = v;
    if (norm) {
        float4 n;
        n.x = -amp * sinf(ax) * freq;
        n.z = amp * cosf(ay) * freq;
        n.y = 1.0f - n.x * n.x - n.z * n.z;
        n.w = 0.0f;
        norm[id] = n;
    }
}

// radial wave on unit-sphere vertices: r = 1 + A sin(kt theta + wt frame)
// cos(kp phi + wp frame) with theta = acos(z), phi = atan2(y, x)
__global__ __launch_bounds__(TB) void radial_wave_kernel(const float4 *__restrict__ unit, int nv,
                                                         float frame, float amp, float kt, float wt,
                                                         float kp, float wp,
                                                         float4 *__restrict__ out) {
    const int id = blockIdx.x * TB + threadIdx.x;
    if (id >= nv) return;
    const float4 u = unit[id];
    const float theta = acosf(fminf(fmaxf(u.z, -1.0f), 1.0f));
    const float phi = atan2f(u.y, u.x);
    const float r = 1.0f + amp * sinf(kt * theta + wt * frame) * cosf(kp * phi + wp * frame);
    out[id] = make_float4(u.x * r, u.y * r, u.z * r, u.w);
}

// AddMeshKernel's normals and colours (GeometryList.cu:215-230):
// n_k = (normalMat * normal.xyz, 0), normalMat rows = model rows' xyz
struct Mat3 { float4 r[3]; };

__global__ __launch_bounds__(TB) void add_attributes_kernel(const float4 *__restrict__ normals,
                                                            const uchar4 *__restrict__ colors, int nv,
                                                            const unsigned *__restrict__ idx, int ntris,
                                                            Mat3 m, float4 *__restrict__ n0,
                                                            float4 *__restrict__ n1, float4 *__restrict__ n2,
                                                            uchar4 *__restrict__ c0, uchar4 *__restrict__ c1,
                                                            uchar4 *__restrict__ c2, int *bad) {
    const int id = blockIdx.x * TB + threadIdx.x;
    if (id >= ntris) return;
    const unsigned i[3] = {idx[3 * id], idx[3 * id + 1], idx[3 * id + 2]};
    if (i[0] >= (unsigned)nv || i[1] >= (unsigned)nv || i[2] >= (unsigned)nv) {
        atomicOr(bad, 1);
        return;
    }
    float4 *nout[3] = {n0, n1, n2};
    uchar4 *cout[3] = {c0, c1, c2};
    for (int k = 0; k < 3; ++k) {
        const float4 v = normals[i[k]];
        const f3 n = {v.x, v.y, v.z};
        nout[k][id] = make_float4(dot3(mk3(m.r[0]), n), dot3(mk3(m.r[1]), n), dot3(mk3(m.r[2]), n), 0.0f);
        if (colors) cout[k][id] = colors[i[k]];
    }
}

__global__ void clear_flag_kernel(int *flag) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *flag = 0;
}

}  // namespace

void add_mesh_attributes(pm_ctx *ctx, const float4 *normals, const uchar4 *colors, int nv, const unsigned *idx,
                         int nidx, const float *model, float4 *n0, float4 *n1, float4 *n2, uchar4 *c0,
                         uchar4 *c1, uchar4 *c2, hipStream_t s) {
    Mat3 m;
    for (int i = 0; i < 3; ++i) {
        if (model)
            m.r[i] = make_float4(model[4 * i], model[4 * i + 1], model[4 * i + 2], 0.0f);
        else
            m.r[i] = make_float4(i == 0, i == 1, i == 2, 0.0f);
    }
    const int ntris = nidx / 3;
    BuildScratch &sc = ctx->scratch;
    sc.totals.reserve(8, s);
    sc.host.ensure(8);
    clear_flag_kernel<<<1, 64, 0, s>>>(sc.totals.ptr);
    PM_LAUNCH_CHECK();
    if (ntris > 0) {
        add_attributes_kernel<<<(ntris + TB - 1) / TB, TB, 0, s>>>(normals, colors, nv, idx, ntris, m, n0, n1, n2,
                                                                   c0, c1, c2, sc.totals.ptr);
        PM_LAUNCH_CHECK();
    }
    PM_HIP(hipMemcpyAsync(sc.host.ptr, sc.totals.ptr, sizeof(int), hipMemcpyDeviceToHost, s));
    PM_HIP(hipStreamSynchronize(s));
    if (sc.host.ptr[0]) throw Error{PM_ERR_ARG, "mesh index out of range of the vertex array"};
}

void add_mesh(pm_ctx *ctx, const float4 *verts, int nv, const unsigned *idx, int nidx,
              const float *model, float4 *p0, float4 *p1, float4 *p2, hipStream_t s) {
    Mat4 m;
    for (int i = 0; i < 4; ++i) {
        if (model)
            m.r[i] = make_float4(model[4 * i], model[4 * i + 1], model[4 * i + 2], model[4 * i + 3]);
        else
            m.r[i] = make_float4(i == 0, i == 1, i == 2, i == 3);
    }
    const int ntris = nidx / 3;
    BuildScratch &sc = ctx->scratch;
    sc.totals.reserve(8, s);
    sc.host.ensure(8);
    clear_flag_kernel<<<1, 64, 0, s>>>(sc.totals.ptr);
    PM_LAUNCH_CHECK();
    if (ntris > 0) {
        add_mesh_kernel<<<(ntris + TB - 1) / TB, TB, 0, s>>>(verts, nv, idx, ntris, m, p0, p1, p2, sc.totals.ptr);
        PM_LAUNCH_CHECK();
    }
    PM_HIP(hipMemcpyAsync(sc.host.ptr, sc.totals.ptr, sizeof(int), hipMemcpyDeviceToHost, s));
    PM_HIP(hipStreamSynchronize(s));
    if (sc.host.ptr[0]) throw Error{PM_ERR_ARG, "mesh index out of range of the vertex array"};
}

void osc_surface(float4 *vert, float4 *norm, int d, float height, float time, float amp, float freq,
                 hipStream_t s) {
    const long long n = (long long)d * d;
    if (n == 0) return;
    osc_surface_kernel<<<(int)((n + TB - 1) / TB), TB, 0, s>>>(vert, norm, d, height, time, amp, freq);
    PM_LAUNCH_CHECK();
}

void radial_wave(const float4 *unit, int nv, float frame, const float *w, float4 *out, hipStream_t s) {
    if (nv == 0) return;
    radial_wave_kernel<<<(nv + TB - 1) / TB, TB, 0, s>>>(unit, nv, frame, w[0], w[1], w[2], w[3], w[4], out);
    PM_LAUNCH_CHECK();
}

}  // namespace pmkd
