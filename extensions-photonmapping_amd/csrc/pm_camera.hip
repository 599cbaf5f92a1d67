// pm_camera.hip -- camera rays: IRayTracer::CreateInitialRays / CreateRays /
// Unproject (IRayTracer.cu:45-100) on the device.
//
// Ray id = y * width + x (x = id % width, y = id / width).  The screen
// position is s = ((x + jx) / width) * 2 - 1 (and the same for y with
// height), with (jx, jy) = 0 as in the reference, or a seeded sub-pixel
// jitter (PCG32 stream (seed, purpose 1, id), two 24-bit uniforms).  The
// inverse view-projection m (16 floats, indexed as Unproject indexes them)
// maps (s.x, s.y, 1, 1) to
//   c_k = m[k] s.x + m[4+k] s.y + m[8+k] + m[12+k]   (k = 0..3, left to right),
// the ray's far point is c.xyz / c.w, and the direction is that point minus
// the camera position, normalised as v * (1 / sqrt(v.v)) (the host branch of
// cutil_math's normalize).  Origins are (camera, 1), directions (d, 0).  All
// arithmetic is IEEE binary32 in this order (no contraction), so the rays
// are bit-identical to the CPU restatement of the same steps.
#include "pm_device.h"
#include "pm_internal.h"

namespace pmkd {

namespace {

constexpr int TB = 256;

struct CamArgs {
    float m[16];
    float cx, cy, cz;
};

__global__ __launch_bounds__(TB) void create_rays_kernel(CamArgs cam, int width, int height, int n,
                                                         unsigned long long seed, int jitter,
                                                         float4 *__restrict__ org, float4 *__restrict__ dir) {
    const int id = blockIdx.x * TB + threadIdx.x;
    if (id >= n) return;
    const int x = id % width, y = id / width;
    float jx = 0.0f, jy = 0.0f;
    if (jitter) {
        Pcg rng(seed, (1ull << 40) | (unsigned long long)id);
        jx = rng.uniform();
        jy = rng.uniform();
    }
    const float sx = (((float)x + jx) / (float)width) * 2.0f - 1.0f;
    const float sy = (((float)y + jy) / (float)height) * 2.0f - 1.0f;
    const float *m = cam.m;
    const float c0 = m[0] * sx + m[4] * sy + m[8] + m[12];
    const float c1 = m[1] * sx + m[5] * sy + m[9] + m[13];
    const float c2 = m[2] * sx + m[6] * sy + m[10] + m[14];
    const float c3 = m[3] * sx + m[7] * sy + m[11] + m[15];
    const f3 end = {c0 / c3, c1 / c3, c2 / c3};
    const f3 d = normalize3(sub3(end, f3{cam.cx, cam.cy, cam.cz}));
    org[id] = make_float4(cam.cx, cam.cy, cam.cz, 1.0f);
    dir[id] = make_float4(d.x, d.y, d.z, 0.0f);
}

}  // namespace

void create_rays(const float *inv_view_proj, const float *cam_pos, int width, int height,
                 unsigned long long seed, int jitter, float4 *org, float4 *dir, hipStream_t s) {
    const long long n = (long long)width * height;
    if (n == 0) return;
    CamArgs a;
    for (int k = 0; k < 16; ++k) a.m[k] = inv_view_proj[k];
    a.cx = cam_pos[0]; a.cy = cam_pos[1]; a.cz = cam_pos[2];
    create_rays_kernel<<<(int)((n + TB - 1) / TB), TB, 0, s>>>(a, width, height, (int)n, seed, jitter, org, dir);
    PM_LAUNCH_CHECK();
}

}  // namespace pmkd
