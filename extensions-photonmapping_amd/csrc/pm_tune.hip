// pm_tune.hip -- autotuner: times the caller's workload over the build's
// chunk size, the upper stage's leaf threshold, the short-stack depth and
// the ray order, and keeps the fastest setting.  Coordinate descent, one parameter at a
// time, each setting timed with HIP events on the caller's stream (median of
// `reps` runs after one warm-up run).  The build's own host read-backs are
// inside the timed span, so build times are what a caller waits for.
#include <algorithm>
#include <vector>

#include "pm_internal.h"

namespace pmkd {

namespace {

struct Events {
    hipEvent_t a = nullptr, b = nullptr;
    Events() {
        PM_HIP(hipEventCreate(&a));
        PM_HIP(hipEventCreate(&b));
    }
    ~Events() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
    float ms() const {
        PM_HIP(hipEventSynchronize(b));
        float t = 0.0f;
        PM_HIP(hipEventElapsedTime(&t, a, b));
        return t;
    }
};

float median(std::vector<float> v) {
    std::sort(v.begin(), v.end());
    const size_t n = v.size();
    return n % 2 ? v[n / 2] : 0.5f * (v[n / 2 - 1] + v[n / 2]);
}

}  // namespace

void autotune(pm_ctx *ctx, const float4 *P0, const float4 *P1, const float4 *P2, int n,
              const int *mesh_offsets, int M, const float4 *org, const float4 *dir,
              const int *ray_mesh, int nrays, const pm_build_params &base, float builds_per_trace,
              int reps, pm_tuning &out, hipStream_t s) {
    static const int kChunks[] = {256, 512, 1024, 2048};
    static const int kLeaf[] = {8, 16, 24, 32, 48, 64};
    static const int kStack[] = {4, 8, 16};
    static const int kCap[] = {0, 4};
    DevBuf<pm_hit> hits;
    hits.reserve((size_t)std::max(nrays, 1), s);
    Events ev;
    int trials = 0;
    pm_trace_params tp;
    pm_trace_params_default(&tp);

    auto time_build = [&](const pm_build_params &p) {
        std::vector<float> t;
        for (int r = 0; r <= reps; ++r) {
            PM_HIP(hipEventRecord(ev.a, s));
            build_tree(ctx, P0, P1, P2, n, mesh_offsets, M, p, s);
            PM_HIP(hipEventRecord(ev.b, s));
            const float ms = ev.ms();
            if (r) t.push_back(ms);
        }
        return median(t);
    };
    auto time_trace = [&](int stack, int order, int cap) {
        pm_trace_params q = tp;
        q.stack_size = stack;
        q.ray_order = order;
        q.step_cap = cap;
        std::vector<float> t;
        for (int r = 0; r <= reps; ++r) {
            PM_HIP(hipEventRecord(ev.a, s));
            trace_rays(ctx, PM_TRACE_SHORT_STACK, org, dir, nrays, ray_mesh, hits.ptr, q, nullptr, s);
            PM_HIP(hipEventRecord(ev.b, s));
            const float ms = ev.ms();
            if (r) t.push_back(ms);
        }
        return median(t);
    };

    pm_build_params p = base;
    // 1. chunk size: build time only (the tree does not depend on it; a build
    // that fails here fails for every chunk size and is reported)
    float best = 0.0f;
    int best_chunk = 1024;
    for (int c : kChunks) {
        p.chunk_size = c;
        const float t = time_build(p);
        ++trials;
        if (c == kChunks[0] || t < best) { best = t; best_chunk = c; }
    }
    p.chunk_size = best_chunk;
    // 2. leaf threshold: builds_per_trace * build + trace at the default depth.
    // A threshold the geometry cannot reach (more than that many triangles
    // share a point, so the reference-count guard stops the build) is skipped.
    int best_leaf = base.lower_size;
    float best_cost = 0.0f, best_b = 0.0f;
    bool first = true;
    for (int ls : kLeaf) {
        p.lower_size = ls;
        float b;
        try {
            b = time_build(p);
        } catch (const Error &e) {
            if (e.code != PM_ERR_DEPTH) throw;
            continue;
        }
        const float t = time_trace(8, PM_RAY_ORDER_INPUT, 0);
        ++trials;
        const float cost = builds_per_trace * b + t;
        if (first || cost < best_cost) { best_cost = cost; best_leaf = ls; best_b = b; first = false; }
    }
    if (first) throw Error{PM_ERR_DEPTH, "no leaf threshold builds this geometry"};
    p.lower_size = best_leaf;
    // 3. stack depth, ray order and step cap on the chosen tree (the order
    // only matters from the size at which the tracers apply it)
    build_tree(ctx, P0, P1, P2, n, mesh_offsets, M, p, s);
    int best_stack = 8, best_order = PM_RAY_ORDER_INPUT, best_cap = 0;
    float best_t = 0.0f;
    first = true;
    for (int order : {PM_RAY_ORDER_INPUT, PM_RAY_ORDER_COHERENT}) {
        pm_trace_params q = tp;
        q.ray_order = order;
        if (order != PM_RAY_ORDER_INPUT && !ray_order_applies(q, nrays)) continue;
        for (int S : kStack)
            for (int cap : kCap) {
                const float t = time_trace(S, order, cap);
                ++trials;
                if (first || t < best_t) {
                    best_t = t; best_stack = S; best_order = order; best_cap = cap; first = false;
                }
            }
    }
    PM_HIP(hipStreamSynchronize(s));
    out.chunk_size = best_chunk;
    out.lower_size = best_leaf;
    out.stack_size = best_stack;
    out.build_ms = best_b;
    out.trace_ms = best_t;
    out.trials = trials;
    out.ray_order = best_order;
    out.step_cap = best_cap;
}

}  // namespace pmkd
