// IRayTracer.hpp -- tracer interface (Utils/CUDA/IRayTracer.h:30-60).
// Trace() takes device ray arrays (float4 origin/direction) and writes the
// first hit of each ray; the reference builds camera rays itself and shades
// into a canvas (IRayTracer.cu:72-100, Raytracer.cu:295-390).
#pragma once
#include "TriangleMap.hpp"

namespace pmkd {
namespace Utils {

class IRayTracer {
public:
    enum IntersectionAlgorithm { MOELLER, WOOP };

    IRayTracer() {
        pm_trace_params_default(&params_);
        hip_ok(hipEventCreate(&e0_), "hipEventCreate");
        hip_ok(hipEventCreate(&e1_), "hipEventCreate");
    }
    virtual ~IRayTracer() {
        (void)hipEventDestroy(e0_);
        (void)hipEventDestroy(e1_);
    }
    IRayTracer(const IRayTracer &) = delete;
    IRayTracer &operator=(const IRayTracer &) = delete;

    void Trace(const float4 *origins, const float4 *directions, int rays, pm_hit *hits,
               hipStream_t stream = nullptr, int *counters = nullptr) {
        if (algorithm_ == WOOP) throw NotImplemented("Woop intersection is not provided");
        hip_ok(hipEventRecord(e0_, stream), "hipEventRecord");
        check(Ctx(), pm_trace(Ctx(), Kind(), (const float *)origins, (const float *)directions, rays, hits,
                              &params_, counters, stream));
        hip_ok(hipEventRecord(e1_, stream), "hipEventRecord");
        if (printTiming_) {
            hip_ok(hipEventSynchronize(e1_), "hipEventSynchronize");
            hip_ok(hipEventElapsedTime(&renderTime_, e0_, e1_), "hipEventElapsedTime");
        }
    }
    // HostTrace (IRayTracer.h:48): one ray, result returned to the host
    pm_hit HostTrace(const float4 &origin, const float4 &direction) {
        DeviceBuffer<float4> o(1), d(1);
        DeviceBuffer<pm_hit> h(1);
        o.Upload({origin});
        d.Upload({direction});
        Trace(o.GetDeviceData(), d.GetDeviceData(), 1, h.GetDeviceData());
        return h.Download()[0];
    }

    void SetIntersectionAlgorithm(IntersectionAlgorithm a) { algorithm_ = a; }
    IntersectionAlgorithm GetIntersectionAlgorithm() const { return algorithm_; }
    void SetLeafSkipping(bool s) { params_.leaf_skipping = s; }
    bool GetLeafSkipping() const { return params_.leaf_skipping != 0; }
    void PrintTiming(bool p) { printTiming_ = p; }
    bool GetPrintTiming() const { return printTiming_; }
    // trace time of the last Trace() in ms (with PrintTiming(true))
    float GetRenderTime() const { return renderTime_; }
    pm_trace_params &Params() { return params_; }

protected:
    virtual pm_ctx *Ctx() const = 0;
    virtual int Kind() const = 0;
    pm_trace_params params_;

private:
    IntersectionAlgorithm algorithm_ = MOELLER;
    bool printTiming_ = false;
    float renderTime_ = 0.0f;
    hipEvent_t e0_ = nullptr, e1_ = nullptr;
};

}  // namespace Utils
}  // namespace pmkd
