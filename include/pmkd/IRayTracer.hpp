// IRayTracer.hpp -- tracer interface (Utils/CUDA/IRayTracer.h:30-60).
// Trace() takes a Camera (the rays are made on the device as
// CreateInitialRays makes them, IRayTracer.cu:72-100) or device ray arrays
// (float4 origin/direction), and writes the first hit of each ray, or (with
// a Shading) the shaded canvas the reference renders (Raytracer.cu:165-262,
// ShortStack.cu:96-194).
#pragma once
#include "TriangleMap.hpp"

namespace pmkd {
namespace Utils {

class IRayTracer {
public:
    enum IntersectionAlgorithm { MOELLER = PM_INTERSECT_MOELLER, WOOP = PM_INTERSECT_WOOP };

    // what the canvas shading reads (GeometryList n0..n2/c0, ColorKernels.h:10-13)
    struct Shading {
        const float4 *n0, *n1, *n2;
        const uchar4 *c0;
        pm_light light;
        int max_bounces = 64;
    };

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
        params_.intersection = algorithm_;
        hip_ok(hipEventRecord(e0_, stream), "hipEventRecord");
        check(Ctx(), pm_trace(Ctx(), Kind(), (const float *)origins, (const float *)directions, rays, hits,
                              &params_, counters, stream));
        hip_ok(hipEventRecord(e1_, stream), "hipEventRecord");
        if (printTiming_) {
            hip_ok(hipEventSynchronize(e1_), "hipEventSynchronize");
            hip_ok(hipEventElapsedTime(&renderTime_, e0_, e1_), "hipEventElapsedTime");
        }
    }
    // Trace(canvas) (IRayTracer.h:47): one shaded uchar4 per ray
    void Trace(const float4 *origins, const float4 *directions, int rays, const Shading &sh, uchar4 *canvas,
               hipStream_t stream = nullptr) {
        params_.intersection = algorithm_;
        hip_ok(hipEventRecord(e0_, stream), "hipEventRecord");
        check(Ctx(), pm_trace_canvas(Ctx(), Kind(), (const float *)origins, (const float *)directions, rays,
                                     (const float *)sh.n0, (const float *)sh.n1, (const float *)sh.n2,
                                     (const unsigned char *)sh.c0, &sh.light, sh.max_bounces,
                                     (unsigned char *)canvas, &params_, stream));
        hip_ok(hipEventRecord(e1_, stream), "hipEventRecord");
        if (printTiming_) {
            hip_ok(hipEventSynchronize(e1_), "hipEventSynchronize");
            hip_ok(hipEventElapsedTime(&renderTime_, e0_, e1_), "hipEventElapsedTime");
        }
    }
    // The camera CreateInitialRays reads from the canvas's viewing volume
    // (IRayTracer.cu:72-100): the inverse of view * projection as 16 floats
    // in the element order Unproject reads them (IRayTracer.cu:45-53), the
    // camera position and the canvas size; jitter adds a seeded sub-pixel
    // offset (off in the reference).
    struct Camera {
        float inv_view_proj[16];
        float position[3];
        int width = 0, height = 0;
        bool jitter = false;
        unsigned long long jitter_seed = 0;
    };
    // CreateInitialRays (IRayTracer.cu:72-100): one ray per pixel into the
    // tracer's own origin/direction buffers (GetOrigins / GetDirections)
    void CreateInitialRays(const Camera &cam, hipStream_t stream = nullptr) {
        const size_t n = (size_t)cam.width * (size_t)cam.height;
        origin_.Extend(n);
        direction_.Extend(n);
        check(Ctx(), pm_create_rays(Ctx(), cam.inv_view_proj, cam.position, cam.width, cam.height,
                                    cam.jitter_seed, cam.jitter ? 1 : 0, (float *)origin_.GetDeviceData(),
                                    (float *)direction_.GetDeviceData(), stream));
    }
    const float4 *GetOrigins() const { return origin_.GetDeviceData(); }
    const float4 *GetDirections() const { return direction_.GetDeviceData(); }
    // Trace(IRenderCanvas*, uchar4*) (IRayTracer.h:47): camera rays, then one
    // shaded uchar4 per pixel (id = y * width + x)
    void Trace(const Camera &cam, const Shading &sh, uchar4 *canvas, hipStream_t stream = nullptr) {
        CreateInitialRays(cam, stream);
        Trace(origin_.GetDeviceData(), direction_.GetDeviceData(), cam.width * cam.height, sh, canvas, stream);
    }
    // ... and the first hit of every camera ray
    void Trace(const Camera &cam, pm_hit *hits, hipStream_t stream = nullptr) {
        CreateInitialRays(cam, stream);
        Trace(origin_.GetDeviceData(), direction_.GetDeviceData(), cam.width * cam.height, hits, stream);
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
    // no reference counterpart: large traces of incoherent rays may run in
    // a coherent ray order (off by default); the hit records are the same
    void SetCoherentRayOrder(bool on) { params_.ray_order = on ? PM_RAY_ORDER_COHERENT : PM_RAY_ORDER_INPUT; }
    // no reference counterpart: inner nodes a descent walks per round of the
    // wave-cooperative leaf test (0 = no cap); the hit records are the same
    void SetStepCap(int cap) { params_.step_cap = cap; }
    void PrintTiming(bool p) { printTiming_ = p; }
    bool GetPrintTiming() const { return printTiming_; }
    // trace time of the last Trace() in ms (with PrintTiming(true))
    float GetRenderTime() const { return renderTime_; }
    pm_trace_params &Params() { return params_; }
    // the context the tracer runs in (its tree and triangles)
    pm_ctx *Context() const { return Ctx(); }

protected:
    virtual pm_ctx *Ctx() const = 0;
    virtual int Kind() const = 0;
    pm_trace_params params_;

private:
    IntersectionAlgorithm algorithm_ = MOELLER;
    DeviceBuffer<float4> origin_, direction_;   // camera rays (IRayTracer::origin/direction)
    bool printTiming_ = false;
    float renderTime_ = 0.0f;
    hipEvent_t e0_ = nullptr, e1_ = nullptr;
};

}  // namespace Utils
}  // namespace pmkd
