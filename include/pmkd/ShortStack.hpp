// ShortStack.hpp -- short-stack tracer over a TriangleMap
// (Utils/CUDA/ShortStack.h:25-111, ShortStack.cu:58-194).  The stack depth
// is SHORT_STACK_SIZE (ShortStack.cu:25, 4) by default and settable 1..16.
#pragma once
#include "IRayTracer.hpp"

namespace pmkd {
namespace Utils {

class ShortStack : public IRayTracer {
public:
    explicit ShortStack(TriangleMap *map) : map_(map) {}
    void SetStackSize(int n) { params_.stack_size = n; }
    int GetStackSize() const { return params_.stack_size; }

protected:
    pm_ctx *Ctx() const override { return map_->Context(); }
    int Kind() const override { return PM_TRACE_SHORT_STACK; }

private:
    TriangleMap *map_;
};

}  // namespace Utils
}  // namespace pmkd
