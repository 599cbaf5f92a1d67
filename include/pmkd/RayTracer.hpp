// RayTracer.hpp -- kd-restart tracer over a TriangleMap
// (Utils/CUDA/Raytracer.h:26-44, Raytracer.cu:57-262).
#pragma once
#include "IRayTracer.hpp"

namespace pmkd {
namespace Utils {

class RayTracer : public IRayTracer {
public:
    explicit RayTracer(TriangleMap *map) : map_(map) {}

protected:
    pm_ctx *Ctx() const override { return map_->Context(); }
    int Kind() const override { return PM_TRACE_KD_RESTART; }

private:
    TriangleMap *map_;
};

}  // namespace Utils
}  // namespace pmkd
