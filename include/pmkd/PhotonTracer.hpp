// PhotonTracer.hpp -- photon emission and bounce driver over a tracer.  No
// counterpart in the reference: its tracers serve a photon-mapping renderer
// the extension does not contain.  Emits photons from area lights
// (pm_photon_emit), traces them with the given IRayTracer (its algorithm and
// options), and continues every photon that hit a triangle below
// absorb_from with a cosine-weighted bounce ray (pm_photon_bounce), up to
// `bounces` times.  Each pass's rays, hits and source indices stay on the
// device.  The ray order (coherent sort or input order) is chosen per pass:
// SetRayOrders, or TuneRayOrders, which times both orders of every pass and
// keeps the faster one.  The Python driver photons.PhotonTracer is the same
// loop.
#pragma once
#include <algorithm>
#include <vector>

#include "IRayTracer.hpp"

namespace pmkd {
namespace Utils {

class PhotonTracer {
public:
    // lights: device light triangles (l0/l1/l2 float4[lights]); photons that
    // reach a triangle id >= absorb_from end there
    PhotonTracer(IRayTracer &tracer, const float4 *l0, const float4 *l1, const float4 *l2, int lights,
                 int absorb_from)
        : tracer_(tracer), l0_(l0), l1_(l1), l2_(l2), lights_(lights), absorb_from_(absorb_from) {
        hip_ok(hipEventCreate(&e_[0]), "hipEventCreate");
        hip_ok(hipEventCreate(&e_[1]), "hipEventCreate");
    }
    ~PhotonTracer() {
        (void)hipEventDestroy(e_[0]);
        (void)hipEventDestroy(e_[1]);
    }
    PhotonTracer(const PhotonTracer &) = delete;
    PhotonTracer &operator=(const PhotonTracer &) = delete;

    void SetOffset(float offset) { offset_ = offset; }
    // the ray order of each pass (PM_RAY_ORDER_*); later passes use the last
    void SetRayOrders(const std::vector<int> &orders) {
        if (orders.empty()) throw Exception(PM_ERR_ARG, "at least one ray order");
        for (int o : orders)
            if (o != PM_RAY_ORDER_INPUT && o != PM_RAY_ORDER_COHERENT) throw Exception(PM_ERR_ARG, "unknown ray order");
        orders_ = orders;
    }
    const std::vector<int> &GetRayOrders() const { return orders_; }
    int RayOrderOf(int pass) const { return orders_[std::min<size_t>(pass, orders_.size() - 1)]; }

    // emission + up to `bounces` bounce passes (pass b + 1 uses the bounce
    // seed seed + 1000 (b + 1)); returns the rays of each pass traced.  With
    // timed, GetPassTimes() holds each pass's trace time in ms.
    const std::vector<int> &Trace(int photons, int bounces, unsigned long long seed, hipStream_t stream = nullptr,
                                  bool timed = false) {
        if (photons < 0 || bounces < 0) throw Exception(PM_ERR_ARG, "photons and bounces must be >= 0");
        Reserve(photons, bounces + 1);
        pm_ctx *ctx = tracer_.Context();
        check(ctx, pm_photon_emit(ctx, (const float *)l0_, (const float *)l1_, (const float *)l2_, lights_, photons,
                                  seed, offset_, (float *)pass_[0].org.GetDeviceData(),
                                  (float *)pass_[0].dir.GetDeviceData(), stream));
        rays_.clear();
        times_.clear();
        const int saved = tracer_.Params().ray_order;
        int n = photons;
        for (int b = 0; b <= bounces; ++b) {
            Pass &p = pass_[b];
            tracer_.Params().ray_order = RayOrderOf(b);
            if (timed) hip_ok(hipEventRecord(e_[0], stream), "hipEventRecord");
            if (n > 0) tracer_.Trace(p.org.GetDeviceData(), p.dir.GetDeviceData(), n, p.hits.GetDeviceData(), stream);
            if (timed) {
                hip_ok(hipEventRecord(e_[1], stream), "hipEventRecord");
                hip_ok(hipEventSynchronize(e_[1]), "hipEventSynchronize");
                float ms = 0.0f;
                hip_ok(hipEventElapsedTime(&ms, e_[0], e_[1]), "hipEventElapsedTime");
                times_.push_back(ms);
            }
            rays_.push_back(n);
            if (b == bounces || n == 0) break;
            Pass &q = pass_[b + 1];
            int m = 0;
            check(ctx, pm_photon_bounce(ctx, (const float *)p.org.GetDeviceData(), (const float *)p.dir.GetDeviceData(),
                                        p.hits.GetDeviceData(), n, absorb_from_, seed + 1000ull * (b + 1), offset_,
                                        (float *)q.org.GetDeviceData(), (float *)q.dir.GetDeviceData(),
                                        q.src.GetDeviceData(), nullptr, &m, stream));
            n = m;
        }
        tracer_.Params().ray_order = saved;
        return rays_;
    }

    // time every pass in both orders (after one untimed run of each) and
    // keep the faster order per pass; the hit records are the same either way
    const std::vector<int> &TuneRayOrders(int photons, int bounces, unsigned long long seed,
                                          hipStream_t stream = nullptr) {
        std::vector<float> t[2];
        const int order[2] = {PM_RAY_ORDER_COHERENT, PM_RAY_ORDER_INPUT};
        for (int k = 0; k < 2; ++k) {
            orders_ = {order[k]};
            Trace(photons, bounces, seed, stream);
            Trace(photons, bounces, seed, stream, true);
            t[k] = times_;
        }
        std::vector<int> best;
        for (size_t b = 0; b < t[0].size(); ++b)
            best.push_back(b < t[1].size() && t[1][b] < t[0][b] ? PM_RAY_ORDER_INPUT : PM_RAY_ORDER_COHERENT);
        if (best.empty()) best.push_back(PM_RAY_ORDER_COHERENT);
        orders_ = best;
        return orders_;
    }

    const std::vector<int> &GetPassRays() const { return rays_; }
    const std::vector<float> &GetPassTimes() const { return times_; }
    const float4 *GetOrigins(int pass) const { return pass_.at(pass).org.GetDeviceData(); }
    const float4 *GetDirections(int pass) const { return pass_.at(pass).dir.GetDeviceData(); }
    const pm_hit *GetHits(int pass) const { return pass_.at(pass).hits.GetDeviceData(); }
    // pass >= 1: the index of the previous pass's ray each ray continues
    const int *GetSources(int pass) const { return pass_.at(pass).src.GetDeviceData(); }

private:
    struct Pass {
        DeviceBuffer<float4> org, dir;
        DeviceBuffer<pm_hit> hits;
        DeviceBuffer<int> src;
    };
    void Reserve(int n, int passes) {
        const size_t m = (size_t)std::max(n, 1);
        while ((int)pass_.size() < passes) pass_.emplace_back();
        for (Pass &p : pass_) {
            p.org.Extend(m);
            p.dir.Extend(m);
            p.hits.Extend(m);
            p.src.Extend(m);
        }
    }

    IRayTracer &tracer_;
    const float4 *l0_, *l1_, *l2_;
    int lights_, absorb_from_;
    float offset_ = 1e-4f;
    std::vector<int> orders_{PM_RAY_ORDER_COHERENT};
    std::vector<Pass> pass_;
    std::vector<int> rays_;
    std::vector<float> times_;
    hipEvent_t e_[2] = {nullptr, nullptr};
};

}  // namespace Utils
}  // namespace pmkd
