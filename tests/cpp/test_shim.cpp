// Drives the C++ API (include/pmkd/*.hpp, the reference's class names) on
// the GPU: mesh staging with a model matrix, TriangleMap::Create, the three
// tracers, HostTrace, options and error behaviour.  Exit status 0 = pass.
#include <cmath>
#include <cstdio>
#include <vector>

#include "pmkd/pmkd.hpp"

using namespace pmkd;

static int failures = 0;
#define EXPECT(c)                                                        \
    do {                                                                 \
        if (!(c)) { std::printf("FAIL %s:%d %s\n", __FILE__, __LINE__, #c); ++failures; } \
    } while (0)

int main() {
    // mesh 1: an axis-aligned box [-1,1]^3 as 12 triangles (CUDAMeshNode)
    std::vector<float4> bv;
    for (int i = 0; i < 8; ++i) bv.push_back(float4{(i & 1) ? 1.f : -1.f, (i & 2) ? 1.f : -1.f, (i & 4) ? 1.f : -1.f, 1.f});
    const unsigned quads[6][4] = {{0, 1, 3, 2}, {4, 6, 7, 5}, {0, 4, 5, 1}, {2, 3, 7, 6}, {0, 2, 6, 4}, {1, 5, 7, 3}};
    std::vector<unsigned> bi;
    for (auto &q : quads) { bi.insert(bi.end(), {q[0], q[1], q[2], q[0], q[2], q[3]}); }
    Scene::CUDAMeshNode box(bv, bi);
    // mesh 2: a 64x64 grid of small triangles, moved by a model matrix
    std::vector<float4> gv;
    std::vector<unsigned> gi;
    const int G = 64;
    for (int j = 0; j <= G; ++j)
        for (int i = 0; i <= G; ++i)
            gv.push_back(float4{i / float(G) - 0.5f, 0.1f * std::sin(i * 0.3f + j * 0.2f), j / float(G) - 0.5f, 1.f});
    for (int j = 0; j < G; ++j)
        for (int i = 0; i < G; ++i) {
            const unsigned a = j * (G + 1) + i, b = a + 1, c = a + G + 1, d = c + 1;
            gi.insert(gi.end(), {a, b, d, a, d, c});
        }
    Scene::CUDAMeshNode grid(gv, gi);
    const float model[16] = {1, 0, 0, 0.1f, 0, 1, 0, -0.5f, 0, 0, 1, 0, 0, 0, 0, 1};   // translate

    Utils::GeometryList geom;
    geom.AddMesh(&box);
    geom.AddMesh(&grid, model);
    EXPECT(geom.GetSize() == 12 + 2 * G * G);
    // staging check: first grid triangle, vertex 0 = (-0.5 + 0.1, 0 - 0.5, -0.5)
    {
        std::vector<float4> p0(geom.GetSize());
        hip_ok(hipMemcpy(p0.data(), geom.GetP0Data(), p0.size() * sizeof(float4), hipMemcpyDeviceToHost), "copy");
        EXPECT(p0[12].x == -0.5f + 0.1f && p0[12].y == 0.0f - 0.5f && p0[12].z == -0.5f && p0[12].w == 1.0f);
    }

    Utils::TriangleMap map(&geom);
    map.SetLowerAlgorithm(Utils::TriangleMap::SAH);
    map.Create();
    EXPECT(map.GetConstructionTime() > 0.0f);
    Scene::TriangleNode nodes = map.GetNodes();
    EXPECT(nodes.GetSize() > 1 && map.GetPrimitiveIndicesSize() >= geom.GetSize());
    const pm_stats st = map.Stats();
    EXPECT(st.reachable_nodes == st.num_nodes && st.leaves == st.inner_nodes + 1);

    // rays from a pinhole at (0, 0.5, 3) towards the scene
    const int W = 128, H = 128, R = W * H;
    std::vector<float4> org(R), dir(R);
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            const float sx = (x + 0.5f) / W * 2 - 1, sy = (y + 0.5f) / H * 2 - 1;
            float d[3] = {sx * 0.5f, sy * 0.5f - 0.2f, -1.f};
            const float n = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            org[y * W + x] = float4{0.f, 0.5f, 3.f, 1.f};
            dir[y * W + x] = float4{d[0] / n, d[1] / n, d[2] / n, 0.f};
        }
    DeviceBuffer<float4> O, D;
    O.Upload(org);
    D.Upload(dir);
    DeviceBuffer<pm_hit> hs(R), hk(R), hb(R);
    Utils::ShortStack ss(&map);
    Utils::RayTracer kd(&map);
    Utils::BruteTracer brute(&geom);
    ss.SetStackSize(8);
    ss.PrintTiming(true);
    ss.Trace(O.GetDeviceData(), D.GetDeviceData(), R, hs.GetDeviceData());
    kd.Trace(O.GetDeviceData(), D.GetDeviceData(), R, hk.GetDeviceData());
    brute.Trace(O.GetDeviceData(), D.GetDeviceData(), R, hb.GetDeviceData());
    hip_ok(hipDeviceSynchronize(), "sync");
    EXPECT(ss.GetRenderTime() > 0.0f);
    const auto a = hs.Download(), b = hk.Download(), c = hb.Download();
    int hits = 0, diff_kd = 0, diff_brute = 0;
    for (int i = 0; i < R; ++i) {
        hits += a[i].id >= 0;
        diff_kd += a[i].id != b[i].id;
        diff_brute += a[i].id != c[i].id;
    }
    EXPECT(hits > R / 2);
    EXPECT(diff_kd <= R / 1000 && diff_brute <= R / 1000);
    const pm_hit one = ss.HostTrace(org[R / 2 + W / 2], dir[R / 2 + W / 2]);
    EXPECT(one.id == a[R / 2 + W / 2].id && one.t == a[R / 2 + W / 2].t);

    // options and errors (reference behaviour: SPLIT throws NotImplemented)
    map.SetSplitMethod(Utils::TriangleMap::SPLIT);
    bool threw = false;
    try { map.Create(); } catch (const NotImplemented &) { threw = true; }
    EXPECT(threw);
    map.SetSplitMethod(Utils::TriangleMap::BOX);
    map.SetLowerAlgorithm(Utils::TriangleMap::BITMAP);
    map.SplitEmptySpace(false);
    map.Create();
    kd.Trace(O.GetDeviceData(), D.GetDeviceData(), R, hk.GetDeviceData());
    hip_ok(hipDeviceSynchronize(), "sync");
    const auto b2 = hk.Download();
    int diff2 = 0;
    for (int i = 0; i < R; ++i) diff2 += b2[i].id != c[i].id;
    EXPECT(diff2 <= R / 1000);
    // Woop intersection (IRayTracer.h:78-109): the same first hits up to rounding
    ss.SetIntersectionAlgorithm(Utils::IRayTracer::WOOP);
    ss.Trace(O.GetDeviceData(), D.GetDeviceData(), R, hs.GetDeviceData());
    hip_ok(hipDeviceSynchronize(), "sync");
    const auto w = hs.Download();
    int diff_woop = 0;
    for (int i = 0; i < R; ++i) diff_woop += w[i].id != b2[i].id;
    EXPECT(diff_woop <= R / 100);
    ss.SetIntersectionAlgorithm(Utils::IRayTracer::MOELLER);

    // the shaded canvas (IRayTracer::Trace(canvas)): every hit pixel is lit
    Utils::IRayTracer::Shading sh{geom.GetNormal0Data(), geom.GetNormal1Data(), geom.GetNormal2Data(),
                                  geom.GetColor0Data(), pm_light{{0.0f, 0.9f, 0.0f}, {0.2f, 0.2f, 0.2f},
                                                                 {0.8f, 0.8f, 0.8f}, {0.0f, 0.0f, 0.0f}}, 8};
    DeviceBuffer<uchar4> canvas(R);
    ss.Trace(O.GetDeviceData(), D.GetDeviceData(), R, sh, canvas.GetDeviceData());
    hip_ok(hipDeviceSynchronize(), "sync");
    const auto px = canvas.Download();
    int lit = 0;
    for (int i = 0; i < R; ++i) lit += (w[i].id >= 0) == (px[i].w > 0);
    EXPECT(lit >= R - R / 100);

    // camera rays made on the device (CreateInitialRays): an affine inverse
    // view-projection whose far point for screen (sx, sy) is
    // (0.5 sx, 0.5 sy + 0.3, 2), seen from (0, 0.5, 3)
    Utils::IRayTracer::Camera cam{};
    const float inv[16] = {0.5f, 0, 0, 0, 0, 0.5f, 0, 0, 0, 0.3f, 2.0f, 0, 0, 0, 0, 1.0f};
    for (int k = 0; k < 16; ++k) cam.inv_view_proj[k] = inv[k];
    cam.position[0] = 0.0f; cam.position[1] = 0.5f; cam.position[2] = 3.0f;
    cam.width = W; cam.height = H;
    DeviceBuffer<pm_hit> hc(R);
    ss.Trace(cam, hc.GetDeviceData());
    hip_ok(hipDeviceSynchronize(), "sync");
    std::vector<float4> cd(R);
    hip_ok(hipMemcpy(cd.data(), ss.GetDirections(), R * sizeof(float4), hipMemcpyDeviceToHost), "copy");
    double worst = 0.0;
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            const double sx = x / double(W) * 2 - 1, sy = y / double(H) * 2 - 1;
            const double v[3] = {0.5 * sx, 0.5 * sy - 0.2, -1.0};
            const double n = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
            const float4 &g = cd[y * W + x];
            worst = std::fmax(worst, std::fabs(g.x - v[0] / n) + std::fabs(g.y - v[1] / n) + std::fabs(g.z - v[2] / n));
        }
    EXPECT(worst < 1e-5);
    const auto ch = hc.Download();
    int cam_hits = 0;
    for (int i = 0; i < R; ++i) cam_hits += ch[i].id >= 0;
    EXPECT(cam_hits > R / 2);
    ss.Trace(cam, sh, canvas.GetDeviceData());
    hip_ok(hipDeviceSynchronize(), "sync");
    const auto cpx = canvas.Download();
    int cam_lit = 0;
    for (int i = 0; i < R; ++i) cam_lit += (ch[i].id >= 0) == (cpx[i].w > 0);
    EXPECT(cam_lit >= R - R / 100);

    // photon driver: the box's y = +1 face (triangles 6, 7, normal -y) emits
    // into the closed box; nothing absorbs, so every pass hits a wall or the
    // grid and the next pass continues all of its rays
    {
        const float4 *P0 = geom.GetP0Data(), *P1 = geom.GetP1Data(), *P2 = geom.GetP2Data();
        Utils::PhotonTracer pt(ss, P0 + 6, P1 + 6, P2 + 6, 2, geom.GetSize());
        const int NP = 20000;
        const std::vector<int> rays = pt.Trace(NP, 2, 77);
        hip_ok(hipDeviceSynchronize(), "sync");
        EXPECT(rays.size() == 3 && rays[0] == NP);
        // each pass's hits are a plain trace of that pass's rays, and the next
        // pass continues every hit (a closed box: all but crack-slipping rays)
        for (int b = 0; b < 3 && b < (int)rays.size(); ++b) {
            DeviceBuffer<pm_hit> hp(rays[b]);
            ss.Trace(pt.GetOrigins(b), pt.GetDirections(b), rays[b], hp.GetDeviceData());
            std::vector<pm_hit> a(rays[b]), w = hp.Download();
            hip_ok(hipMemcpy(a.data(), pt.GetHits(b), rays[b] * sizeof(pm_hit), hipMemcpyDeviceToHost), "copy");
            int same = 0, hit = 0;
            for (int i = 0; i < rays[b]; ++i) {
                same += a[i].id == w[i].id && a[i].t == w[i].t;
                hit += a[i].id >= 0;
            }
            EXPECT(same == rays[b] && hit >= rays[b] - rays[b] / 1000);
            if (b + 1 < (int)rays.size()) EXPECT(rays[b + 1] == hit);
        }
        const std::vector<int> orders = pt.TuneRayOrders(NP, 2, 77);
        EXPECT(orders.size() == 3);
        pt.Trace(NP, 2, 77, nullptr, true);
        EXPECT(pt.GetPassTimes().size() == 3);
    }

    std::printf("%s: %d tris, %d nodes, %d/%d hits, kd diff %d, brute diff %d\n",
                failures ? "FAILED" : "C++ API OK", geom.GetSize(), st.num_nodes, hits, R, diff_kd, diff_brute);
    return failures ? 1 : 0;
}
