// BruteTracer.hpp -- every ray against every triangle
// (Utils/CUDA/BruteTracer.h:26-44, BruteTracer.cu:54-121).  The
// reference's 4096-triangle cap (BruteTracer.cu:50) is off unless set.
#pragma once
#include "IRayTracer.hpp"

namespace pmkd {
namespace Utils {

class BruteTracer : public IRayTracer {
public:
    explicit BruteTracer(GeometryList *geom) : geom_(geom) {
        if (pm_create(&ctx_, 0) != PM_OK) throw Exception(PM_ERR_HIP, "no HIP device");
    }
    ~BruteTracer() override { if (ctx_) pm_destroy(ctx_); }
    void SetPrimitiveCap(int cap) { params_.brute_cap = cap; }

protected:
    pm_ctx *Ctx() const override {
        check(ctx_, pm_set_geometry(ctx_, (const float *)geom_->GetP0Data(), (const float *)geom_->GetP1Data(),
                                    (const float *)geom_->GetP2Data(), geom_->GetSize()));
        return ctx_;
    }
    int Kind() const override { return PM_TRACE_BRUTE; }

private:
    GeometryList *geom_;
    pm_ctx *ctx_ = nullptr;
};

}  // namespace Utils
}  // namespace pmkd
