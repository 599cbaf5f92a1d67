// pm_api.hip -- the C ABI declared in include/pmkd.h.
#include <algorithm>
#include <vector>

#include "pm_internal.h"

using pmkd::Error;

pm_ctx::~pm_ctx() {
    pmkd::destroy_graph(this);
    if (capture_stream) (void)hipStreamDestroy(capture_stream);
}

namespace {

int fail(pm_ctx *ctx, int code, const std::string &msg) {
    if (ctx) ctx->err = msg;
    return code;
}

template <class F>
int guarded(pm_ctx *ctx, F &&f) {
    if (!ctx) return PM_ERR_ARG;
    try {
        PM_HIP(hipSetDevice(ctx->device));
        f();
        ctx->err.clear();
        return PM_OK;
    } catch (const Error &e) {
        return fail(ctx, e.code, e.msg);
    } catch (const std::bad_alloc &) {
        return fail(ctx, PM_ERR_HIP, "host allocation failed");
    }
}

void check_build_params(const pm_build_params &p) {
    if (p.split_method == PM_SPLIT_SPLIT)   // TriangleMapUpperCreator.cu:551-552
        throw Error{PM_ERR_NOT_IMPLEMENTED, "split method SPLIT is not implemented (as in the reference)"};
    if (p.split_method != PM_SPLIT_BOX && p.split_method != PM_SPLIT_DIVIDE) throw Error{PM_ERR_ARG, "unknown split method"};
    if (p.lower_algorithm != PM_LOWER_BITMAP && p.lower_algorithm != PM_LOWER_SAH &&
        p.lower_algorithm != PM_LOWER_BALANCED)
        throw Error{PM_ERR_ARG, "unknown lower algorithm"};
    if (p.lower_size < 1 || p.lower_size > 64) throw Error{PM_ERR_ARG, "lower_size must be in [1, 64]"};
    if (p.max_upper_levels < 0) throw Error{PM_ERR_ARG, "max_upper_levels must be >= 0"};
    if (p.max_ref_factor < 0) throw Error{PM_ERR_ARG, "max_ref_factor must be >= 0"};
    if (p.chunk_size != 0 && p.chunk_size != 128 && p.chunk_size != 256 && p.chunk_size != 512 &&
        p.chunk_size != 1024 && p.chunk_size != 2048 && p.chunk_size != 4096)
        throw Error{PM_ERR_ARG, "chunk_size must be 0, 128, 256, 512, 1024, 2048 or 4096"};
}

void check_trace_params(const pm_trace_params &p) {
    if (p.stack_size < 1 || p.stack_size > 16) throw Error{PM_ERR_ARG, "stack_size must be in [1, 16]"};
    if (p.intersection != PM_INTERSECT_MOELLER && p.intersection != PM_INTERSECT_WOOP)
        throw Error{PM_ERR_ARG, "unknown intersection algorithm"};
    if (p.ray_order != PM_RAY_ORDER_INPUT && p.ray_order != PM_RAY_ORDER_COHERENT)
        throw Error{PM_ERR_ARG, "unknown ray order"};
    if (p.step_cap < 0 || p.step_cap > 4096) throw Error{PM_ERR_ARG, "step_cap must be in [0, 4096]"};
}

}  // namespace

extern "C" {

void pm_build_params_default(pm_build_params *p) {
    if (!p) return;
    p->lower_algorithm = PM_LOWER_BITMAP;
    p->empty_space_splitting = 1;
    p->empty_space_threshold = 0.35f;
    p->split_method = PM_SPLIT_BOX;
    p->propagate_aabbs = 1;
    p->traversal_cost = 24.0f;
    p->lower_size = 32;
    p->max_upper_levels = 0;
    p->sah_leaf_max = 0;
    p->sah_empty_bonus = 1.0f;
    p->max_ref_factor = 0;
    p->chunk_size = 0;
}

void pm_trace_params_default(pm_trace_params *p) {
    if (!p) return;
    p->leaf_skipping = 1;
    p->stack_size = 4;
    p->brute_cap = 0;
    p->intersection = PM_INTERSECT_MOELLER;
    p->ray_order = PM_RAY_ORDER_INPUT;
    p->step_cap = 0;
}

int pm_create(pm_ctx **out, int device) {
    if (!out) return PM_ERR_ARG;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return PM_ERR_HIP;
    if (device < 0 || device >= count) return PM_ERR_ARG;
    pm_ctx *c = new (std::nothrow) pm_ctx();
    if (!c) return PM_ERR_HIP;
    c->device = device;
    *out = c;
    return PM_OK;
}

int pm_destroy(pm_ctx *ctx) {
    if (!ctx) return PM_ERR_ARG;
    (void)hipSetDevice(ctx->device);
    (void)hipDeviceSynchronize();
    delete ctx;
    return PM_OK;
}

const char *pm_last_error(const pm_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int pm_batch_graph_capture(pm_ctx *ctx, const float *p0, const float *p1, const float *p2, int n,
                           const int *mesh_offsets, int num_meshes, const float *org, const float *dir,
                           const int *ray_mesh, int nrays, pm_hit *out, int algo,
                           const pm_build_params *build_params, const pm_trace_params *trace_params,
                           pm_graph_caps *caps, void *stream) {
    return guarded(ctx, [&] {
        if (n < 0 || (n > 0 && (!p0 || !p1 || !p2))) throw Error{PM_ERR_ARG, "invalid geometry"};
        if (num_meshes < 1 || !mesh_offsets) throw Error{PM_ERR_ARG, "need at least one mesh and its offsets"};
        if (nrays < 0 || (nrays > 0 && (!org || !dir || !out || !ray_mesh))) throw Error{PM_ERR_ARG, "invalid rays"};
        if (algo != PM_TRACE_KD_RESTART && algo != PM_TRACE_SHORT_STACK)
            throw Error{PM_ERR_ARG, "a captured batch traces with KD_RESTART or SHORT_STACK"};
        pm_build_params bp;
        if (build_params) bp = *build_params; else pm_build_params_default(&bp);
        check_build_params(bp);
        pm_trace_params tp;
        if (trace_params) tp = *trace_params; else pm_trace_params_default(&tp);
        check_trace_params(tp);
        if (tp.intersection != PM_INTERSECT_MOELLER)
            throw Error{PM_ERR_ARG, "a captured batch traces with MOELLER (Woop matrices are made per call)"};
        pmkd::BuildSizes c;
        if (caps) {
            c.nodes = caps->nodes; c.active = caps->active; c.segs = caps->segments; c.refs = caps->level_refs;
            c.leaf_refs = caps->leaf_refs; c.leaf_ids = caps->leaf_ids; c.lower = caps->lower_nodes;
            c.upper_levels = caps->upper_levels; c.lower_levels = caps->lower_levels;
        }
        pmkd::capture_batch(ctx, reinterpret_cast<const float4 *>(p0), reinterpret_cast<const float4 *>(p1),
                            reinterpret_cast<const float4 *>(p2), n, mesh_offsets, num_meshes,
                            reinterpret_cast<const float4 *>(org), reinterpret_cast<const float4 *>(dir), ray_mesh,
                            nrays, out, algo, bp, tp, c, (hipStream_t)stream);
        if (caps) {
            caps->nodes = c.nodes; caps->active = c.active; caps->segments = c.segs; caps->level_refs = c.refs;
            caps->leaf_refs = c.leaf_refs; caps->leaf_ids = c.leaf_ids; caps->lower_nodes = c.lower;
            caps->upper_levels = c.upper_levels; caps->lower_levels = c.lower_levels;
        }
    });
}

int pm_batch_graph_launch(pm_ctx *ctx, void *stream) {
    return guarded(ctx, [&] { pmkd::launch_batch(ctx, (hipStream_t)stream); });
}

int pm_batch_graph_status(pm_ctx *ctx, int *build_status, void *stream) {
    return guarded(ctx, [&] {
        const int st = pmkd::sync_graph_state(ctx, (hipStream_t)stream);
        if (build_status) *build_status = st;
    });
}

int pm_build(pm_ctx *ctx, const float *p0, const float *p1, const float *p2, int n,
             const pm_build_params *params, void *stream) {
    return guarded(ctx, [&] {
        if (n < 0 || (n > 0 && (!p0 || !p1 || !p2))) throw Error{PM_ERR_ARG, "invalid geometry"};
        pm_build_params prm;
        if (params) prm = *params; else pm_build_params_default(&prm);
        check_build_params(prm);
        pmkd::build_tree(ctx, reinterpret_cast<const float4 *>(p0), reinterpret_cast<const float4 *>(p1),
                         reinterpret_cast<const float4 *>(p2), n, nullptr, 1, prm, (hipStream_t)stream);
    });
}

int pm_build_batched(pm_ctx *ctx, const float *p0, const float *p1, const float *p2, int n,
                     const int *mesh_offsets, int num_meshes, const pm_build_params *params,
                     void *stream) {
    return guarded(ctx, [&] {
        if (n < 0 || (n > 0 && (!p0 || !p1 || !p2))) throw Error{PM_ERR_ARG, "invalid geometry"};
        if (num_meshes < 1 || !mesh_offsets) throw Error{PM_ERR_ARG, "need at least one mesh and its offsets"};
        pm_build_params prm;
        if (params) prm = *params; else pm_build_params_default(&prm);
        check_build_params(prm);
        pmkd::build_tree(ctx, reinterpret_cast<const float4 *>(p0), reinterpret_cast<const float4 *>(p1),
                         reinterpret_cast<const float4 *>(p2), n, mesh_offsets, num_meshes, prm,
                         (hipStream_t)stream);
    });
}

int pm_set_geometry(pm_ctx *ctx, const float *p0, const float *p1, const float *p2, int n) {
    return guarded(ctx, [&] {
        if (n < 0 || (n > 0 && (!p0 || !p1 || !p2))) throw Error{PM_ERR_ARG, "invalid geometry"};
        ctx->p0 = reinterpret_cast<const float4 *>(p0);
        ctx->p1 = reinterpret_cast<const float4 *>(p1);
        ctx->p2 = reinterpret_cast<const float4 *>(p2);
        ctx->ntris = n;
    });
}

int pm_tree_view(pm_ctx *ctx, pm_tree *out) {
    return guarded(ctx, [&] {
        if (!out) throw Error{PM_ERR_ARG, "null output"};
        pmkd::sync_graph_state(ctx, nullptr);
        if (!ctx->has_tree) throw Error{PM_ERR_NO_TREE, "no tree: call pm_build or pm_set_tree first"};
        pm_tree &t = *out;
        t.num_nodes = ctx->num_nodes;
        t.info = ctx->nodes.info.ptr;
        t.split_pos = ctx->nodes.split.ptr;
        t.aabb_min = reinterpret_cast<const float *>(ctx->nodes.bmin.ptr);
        t.aabb_max = reinterpret_cast<const float *>(ctx->nodes.bmax.ptr);
        t.primitive_index = ctx->nodes.prim_index.ptr;
        t.primitive_bitmap = ctx->nodes.bitmap.ptr;
        t.children = reinterpret_cast<const int *>(ctx->nodes.children.ptr);
        t.surface_area = ctx->nodes.area.ptr;
        t.parent = ctx->nodes.parent.ptr;
        t.num_prim_indices = ctx->num_prim_indices;
        t.prim_indices = ctx->prim_indices.ptr;
        t.num_leaf_ids = ctx->num_leaf_ids;
        t.leaf_ids = ctx->leaf_ids.ptr;
        t.p0 = reinterpret_cast<const float *>(ctx->p0);
        t.p1 = reinterpret_cast<const float *>(ctx->p1);
        t.p2 = reinterpret_cast<const float *>(ctx->p2);
        t.num_tris = ctx->ntris;
        t.upper_levels = ctx->upper_levels;
        t.lower_levels = ctx->lower_levels;
        t.has_boxes = ctx->has_boxes ? 1 : 0;
        t.num_roots = ctx->num_roots;
    });
}

int pm_set_tree(pm_ctx *ctx, const pm_tree *h, void *stream) {
    return guarded(ctx, [&] {
        if (!h || h->num_nodes < 1 || !h->info || !h->split_pos || !h->children || !h->primitive_index ||
            !h->primitive_bitmap || h->num_prim_indices < 0 || (h->num_prim_indices > 0 && !h->prim_indices))
            throw Error{PM_ERR_ARG, "invalid host tree"};
        if (h->num_tris < 0 || (h->num_tris > 0 && (!h->p0 || !h->p1 || !h->p2)))
            throw Error{PM_ERR_ARG, "invalid geometry"};
        hipStream_t s = (hipStream_t)stream;
        const size_t n = (size_t)h->num_nodes;
        const int roots = h->num_roots > 0 ? h->num_roots : 1;
        if ((size_t)roots > n) throw Error{PM_ERR_ARG, "more roots than nodes"};
        if (ctx->graph_pending) PM_HIP(hipStreamSynchronize(ctx->graph_last_stream));
        ctx->graph_pending = false;
        ctx->has_tree = false;
        // every index the tracers follow must stay inside its array
        for (size_t i = 0; i < n; ++i) {
            if (h->info[i] & 3) {
                const int l = h->children[2 * i], r = h->children[2 * i + 1];
                if (l < 0 || r < 0 || (size_t)l >= n || (size_t)r >= n)
                    throw Error{PM_ERR_ARG, "child index out of range at node " + std::to_string(i)};
            } else {
                const unsigned long long b = (unsigned long long)h->primitive_bitmap[i];
                if (b && (h->primitive_index[i] < 0 ||
                          (long long)h->primitive_index[i] + 63 - __builtin_clzll(b) >= h->num_prim_indices))
                    throw Error{PM_ERR_ARG, "leaf slots out of range at node " + std::to_string(i)};
            }
        }
        pmkd::NodeArrays &a = ctx->nodes;
        a.reserve(n, s, 0);
        PM_HIP(hipMemcpyAsync(a.info.ptr, h->info, n, hipMemcpyHostToDevice, s));
        PM_HIP(hipMemcpyAsync(a.split.ptr, h->split_pos, n * 4, hipMemcpyHostToDevice, s));
        if (h->aabb_min && h->aabb_max) {
            PM_HIP(hipMemcpyAsync(a.bmin.ptr, h->aabb_min, n * 16, hipMemcpyHostToDevice, s));
            PM_HIP(hipMemcpyAsync(a.bmax.ptr, h->aabb_max, n * 16, hipMemcpyHostToDevice, s));
        }
        PM_HIP(hipMemcpyAsync(a.prim_index.ptr, h->primitive_index, n * 4, hipMemcpyHostToDevice, s));
        PM_HIP(hipMemcpyAsync(a.bitmap.ptr, h->primitive_bitmap, n * 8, hipMemcpyHostToDevice, s));
        PM_HIP(hipMemcpyAsync(a.children.ptr, h->children, n * 8, hipMemcpyHostToDevice, s));
        ctx->prim_indices.reserve((size_t)h->num_prim_indices + 1, s);
        if (h->num_prim_indices > 0)
            PM_HIP(hipMemcpyAsync(ctx->prim_indices.ptr, h->prim_indices, (size_t)h->num_prim_indices * 4,
                                  hipMemcpyHostToDevice, s));
        PM_HIP(hipStreamSynchronize(s));
        ctx->num_nodes = h->num_nodes;
        ctx->num_prim_indices = h->num_prim_indices;
        ctx->num_leaf_ids = 0;
        ctx->upper_levels = h->upper_levels;
        ctx->lower_levels = h->lower_levels;
        ctx->has_boxes = h->aabb_min && h->aabb_max && h->has_boxes;
        ctx->p0 = reinterpret_cast<const float4 *>(h->p0);
        ctx->p1 = reinterpret_cast<const float4 *>(h->p1);
        ctx->p2 = reinterpret_cast<const float4 *>(h->p2);
        ctx->ntris = h->num_tris;
        ctx->num_roots = roots;
        ctx->mesh_offsets.clear();
        pmkd::build_trace_view(ctx, s);
        PM_HIP(hipStreamSynchronize(s));
        ctx->has_tree = true;
    });
}

int pm_tree_stats(pm_ctx *ctx, pm_stats *out, void *stream) {
    return guarded(ctx, [&] {
        if (!out) throw Error{PM_ERR_ARG, "null output"};
        hipStream_t s = (hipStream_t)stream;
        pmkd::sync_graph_state(ctx, s);
        if (!ctx->has_tree) throw Error{PM_ERR_NO_TREE, "no tree"};
        const int n = ctx->num_nodes;
        std::vector<int8_t> info(n);
        std::vector<long long> bmp(n);
        std::vector<int2> ch(n);
        PM_HIP(hipMemcpyAsync(info.data(), ctx->nodes.info.ptr, n, hipMemcpyDeviceToHost, s));
        PM_HIP(hipMemcpyAsync(bmp.data(), ctx->nodes.bitmap.ptr, (size_t)n * 8, hipMemcpyDeviceToHost, s));
        PM_HIP(hipMemcpyAsync(ch.data(), ctx->nodes.children.ptr, (size_t)n * 8, hipMemcpyDeviceToHost, s));
        PM_HIP(hipStreamSynchronize(s));
        pm_stats &st = *out;
        std::memset(&st, 0, sizeof(st));
        st.num_nodes = n;
        st.upper_levels = ctx->upper_levels;
        st.lower_levels = ctx->lower_levels;
        st.num_prim_indices = ctx->num_prim_indices;
        st.num_upper_leaves = ctx->num_leaf_ids;
        st.upper_ref_reads = (long long)ctx->ref_reads;
        st.upper_ref_writes = (long long)ctx->ref_writes;
        st.lower_prims = (long long)ctx->lower_prims;
        st.planned_build = ctx->planned_build ? 1 : 0;
        std::vector<std::pair<int, int>> stk;
        for (int r = ctx->num_roots - 1; r >= 0; --r) stk.push_back({r, 0});
        double depth_sum = 0.0;
        long long guard = 0;
        while (!stk.empty()) {
            if (++guard > 4LL * n + 4) throw Error{PM_ERR_INTERNAL, "tree is not acyclic"};
            const auto [node, depth] = stk.back();
            stk.pop_back();
            if (node < 0 || node >= n) throw Error{PM_ERR_INTERNAL, "child index out of range"};
            st.reachable_nodes++;
            st.max_depth = std::max(st.max_depth, depth);
            if ((info[node] & 3) == 0) {
                st.leaves++;
                const int c = __builtin_popcountll((unsigned long long)bmp[node]);
                if (c == 0) st.empty_leaves++;
                else depth_sum += depth;
                st.leaf_refs += c;
                st.max_leaf_size = std::max(st.max_leaf_size, c);
            } else {
                st.inner_nodes++;
                stk.push_back({ch[node].y, depth + 1});
                stk.push_back({ch[node].x, depth + 1});
            }
        }
        const int nonempty = st.leaves - st.empty_leaves;
        st.mean_leaf_depth = nonempty ? depth_sum / nonempty : 0.0;
    });
}

int pm_trace(pm_ctx *ctx, int algo, const float *org, const float *dir, int nrays, pm_hit *out,
             const pm_trace_params *params, int *counters, void *stream) {
    return guarded(ctx, [&] {
        if (nrays < 0 || (nrays > 0 && (!org || !dir || !out))) throw Error{PM_ERR_ARG, "invalid rays"};
        if (algo != PM_TRACE_BRUTE && algo != PM_TRACE_KD_RESTART && algo != PM_TRACE_SHORT_STACK)
            throw Error{PM_ERR_ARG, "unknown trace algorithm"};
        pm_trace_params prm;
        if (params) prm = *params; else pm_trace_params_default(&prm);
        check_trace_params(prm);
        pmkd::sync_graph_state(ctx, (hipStream_t)stream);   // after a replay: its sizes and status
        if (algo == PM_TRACE_BRUTE) {
            if (ctx->ntris > 0 && !ctx->p0) throw Error{PM_ERR_NO_TREE, "no geometry"};
        } else if (!ctx->has_tree) {
            throw Error{PM_ERR_NO_TREE, "no tree: call pm_build or pm_set_tree first"};
        }
        pmkd::trace_rays(ctx, algo, reinterpret_cast<const float4 *>(org), reinterpret_cast<const float4 *>(dir),
                         nrays, nullptr, out, prm, reinterpret_cast<int4 *>(counters), (hipStream_t)stream);
    });
}

int pm_trace_noalloc(pm_ctx *ctx, int algo, const float *org, const float *dir, int nrays, pm_hit *out,
                     const pm_trace_params *params, int *counters, void *stream) {
    return guarded(ctx, [&] {
        if (nrays < 0 || (nrays > 0 && (!org || !dir || !out))) throw Error{PM_ERR_ARG, "invalid rays"};
        if (algo != PM_TRACE_BRUTE && algo != PM_TRACE_KD_RESTART && algo != PM_TRACE_SHORT_STACK)
            throw Error{PM_ERR_ARG, "unknown trace algorithm"};
        pm_trace_params prm;
        if (params) prm = *params; else pm_trace_params_default(&prm);
        check_trace_params(prm);
        if (prm.intersection != PM_INTERSECT_MOELLER)
            throw Error{PM_ERR_UNSUPPORTED, "pm_trace_noalloc traces with MOELLER only (Woop matrices need memory)"};
        if (pmkd::ray_order_applies(prm, nrays) && nrays > pmkd::NOALLOC_MAX_ORDERED_RAYS)
            throw Error{PM_ERR_ARG, "pm_trace_noalloc orders at most " +
                                        std::to_string(pmkd::NOALLOC_MAX_ORDERED_RAYS) +
                                        " rays coherently (got " + std::to_string(nrays) +
                                        "); use PM_RAY_ORDER_INPUT or pm_trace"};
        pmkd::sync_graph_state(ctx, (hipStream_t)stream);   // after a replay: its sizes and status
        if (algo == PM_TRACE_BRUTE) {
            if (ctx->ntris > 0 && !ctx->p0) throw Error{PM_ERR_NO_TREE, "no geometry"};
        } else if (!ctx->has_tree) {
            throw Error{PM_ERR_NO_TREE, "no tree: call pm_build or pm_set_tree first"};
        }
        const pmkd::StaticScratch &ss = pmkd::static_scratch(ctx->device);
        pmkd::trace_rays(ctx, algo, reinterpret_cast<const float4 *>(org), reinterpret_cast<const float4 *>(dir),
                         nrays, nullptr, out, prm, reinterpret_cast<int4 *>(counters), (hipStream_t)stream, &ss);
    });
}

int pm_trace_batched(pm_ctx *ctx, int algo, const float *org, const float *dir, const int *ray_mesh,
                     int nrays, pm_hit *out, const pm_trace_params *params, int *counters,
                     void *stream) {
    return guarded(ctx, [&] {
        if (nrays < 0 || (nrays > 0 && (!org || !dir || !out || !ray_mesh))) throw Error{PM_ERR_ARG, "invalid rays"};
        if (algo != PM_TRACE_KD_RESTART && algo != PM_TRACE_SHORT_STACK)
            throw Error{PM_ERR_ARG, "batched tracing takes KD_RESTART or SHORT_STACK"};
        pm_trace_params prm;
        if (params) prm = *params; else pm_trace_params_default(&prm);
        check_trace_params(prm);
        pmkd::sync_graph_state(ctx, (hipStream_t)stream);   // after a replay: its sizes and status
        if (!ctx->has_tree) throw Error{PM_ERR_NO_TREE, "no tree: call pm_build_batched first"};
        pmkd::trace_rays(ctx, algo, reinterpret_cast<const float4 *>(org), reinterpret_cast<const float4 *>(dir),
                         nrays, ray_mesh, out, prm, reinterpret_cast<int4 *>(counters), (hipStream_t)stream);
    });
}

int pm_trace_backward(pm_ctx *ctx, const float *p0, const float *p1, const float *p2, int ntris,
                      const float *org, const float *dir, const pm_hit *hits, int nrays,
                      const float *grad_t, const float *grad_u, const float *grad_v, const int *faces,
                      int num_vertices, float *grad_vertices, float *grad_org, float *grad_dir,
                      void *stream) {
    return guarded(ctx, [&] {
        if (nrays < 0 || num_vertices < 0 || ntris < 0 || (nrays > 0 && (!org || !dir || !hits)) ||
            (num_vertices > 0 && !grad_vertices) || (ntris > 0 && (!p0 || !p1 || !p2)))
            throw Error{PM_ERR_ARG, "invalid arguments"};
        // the sort's scan scratch is shared with a captured batch: its last replay must be done
        pmkd::sync_graph_state(ctx, (hipStream_t)stream);
        pmkd::trace_backward(ctx, reinterpret_cast<const float4 *>(p0), reinterpret_cast<const float4 *>(p1),
                             reinterpret_cast<const float4 *>(p2), ntris, reinterpret_cast<const float4 *>(org),
                             reinterpret_cast<const float4 *>(dir), hits, nrays, grad_t, grad_u, grad_v, faces,
                             num_vertices, grad_vertices, grad_org, grad_dir, (hipStream_t)stream);
    });
}

int pm_add_mesh(pm_ctx *ctx, const float *vertices, int nv, const unsigned *indices, int nidx,
                const float *model, float *p0, float *p1, float *p2, int offset, void *stream) {
    return guarded(ctx, [&] {
        if (nv < 0 || nidx < 0 || nidx % 3 || offset < 0 || (nidx > 0 && (!vertices || !indices || !p0 || !p1 || !p2)))
            throw Error{PM_ERR_ARG, "invalid mesh"};
        float4 *q0 = reinterpret_cast<float4 *>(p0) + offset;
        float4 *q1 = reinterpret_cast<float4 *>(p1) + offset;
        float4 *q2 = reinterpret_cast<float4 *>(p2) + offset;
        pmkd::add_mesh(ctx, reinterpret_cast<const float4 *>(vertices), nv, indices, nidx, model, q0, q1, q2,
                       (hipStream_t)stream);
    });
}

int pm_autotune(pm_ctx *ctx, const float *p0, const float *p1, const float *p2, int n,
                const int *mesh_offsets, int num_meshes, const float *org, const float *dir,
                const int *ray_mesh, int nrays, const pm_build_params *base,
                float builds_per_trace, int reps, pm_tuning *out, void *stream) {
    return guarded(ctx, [&] {
        if (n < 0 || (n > 0 && (!p0 || !p1 || !p2))) throw Error{PM_ERR_ARG, "invalid geometry"};
        if (nrays < 1 || !org || !dir) throw Error{PM_ERR_ARG, "the autotuner needs rays"};
        if (!out || reps < 1 || reps > 100 || !(builds_per_trace >= 0.0f))
            throw Error{PM_ERR_ARG, "invalid autotuner arguments"};
        if (mesh_offsets && num_meshes < 1) throw Error{PM_ERR_ARG, "need at least one mesh"};
        if (!mesh_offsets && ray_mesh) throw Error{PM_ERR_ARG, "ray_mesh needs mesh_offsets"};
        pm_build_params prm;
        if (base) prm = *base; else pm_build_params_default(&prm);
        check_build_params(prm);
        pmkd::autotune(ctx, reinterpret_cast<const float4 *>(p0), reinterpret_cast<const float4 *>(p1),
                       reinterpret_cast<const float4 *>(p2), n, mesh_offsets, mesh_offsets ? num_meshes : 1,
                       reinterpret_cast<const float4 *>(org), reinterpret_cast<const float4 *>(dir), ray_mesh,
                       nrays, prm, builds_per_trace, reps, *out, (hipStream_t)stream);
    });
}

int pm_trace_canvas(pm_ctx *ctx, int algo, const float *org, const float *dir, int nrays,
                    const float *n0, const float *n1, const float *n2, const unsigned char *c0,
                    const pm_light *light, int max_bounces, unsigned char *canvas,
                    const pm_trace_params *params, void *stream) {
    return guarded(ctx, [&] {
        if (nrays < 0 || (nrays > 0 && (!org || !dir || !canvas || !n0 || !n1 || !n2 || !c0)) || !light)
            throw Error{PM_ERR_ARG, "invalid arguments"};
        if (max_bounces < 1) throw Error{PM_ERR_ARG, "max_bounces must be >= 1"};
        if (algo != PM_TRACE_BRUTE && algo != PM_TRACE_KD_RESTART && algo != PM_TRACE_SHORT_STACK)
            throw Error{PM_ERR_ARG, "unknown trace algorithm"};
        pm_trace_params prm;
        if (params) prm = *params; else pm_trace_params_default(&prm);
        check_trace_params(prm);
        pmkd::sync_graph_state(ctx, (hipStream_t)stream);
        if (algo == PM_TRACE_BRUTE) {
            if (ctx->ntris > 0 && !ctx->p0) throw Error{PM_ERR_NO_TREE, "no geometry"};
        } else if (!ctx->has_tree) {
            throw Error{PM_ERR_NO_TREE, "no tree: call pm_build or pm_set_tree first"};
        } else if (ctx->num_roots > 1) {
            throw Error{PM_ERR_ARG, "the canvas traces one tree: the context holds a forest of " +
                                        std::to_string(ctx->num_roots) + " meshes"};
        }
        pmkd::trace_canvas(ctx, algo, reinterpret_cast<const float4 *>(org), reinterpret_cast<const float4 *>(dir),
                           nrays, reinterpret_cast<const float4 *>(n0), reinterpret_cast<const float4 *>(n1),
                           reinterpret_cast<const float4 *>(n2), reinterpret_cast<const uchar4 *>(c0), *light,
                           max_bounces, reinterpret_cast<uchar4 *>(canvas), prm, (hipStream_t)stream);
    });
}

int pm_add_mesh_attributes(pm_ctx *ctx, const float *normals, const unsigned char *colors, int nv,
                           const unsigned *indices, int nidx, const float *model, float *n0, float *n1,
                           float *n2, unsigned char *c0, unsigned char *c1, unsigned char *c2, int offset,
                           void *stream) {
    return guarded(ctx, [&] {
        if (nv < 0 || nidx < 0 || nidx % 3 || offset < 0 ||
            (nidx > 0 && (!normals || !indices || !n0 || !n1 || !n2 || (colors && (!c0 || !c1 || !c2)))))
            throw Error{PM_ERR_ARG, "invalid mesh attributes"};
        uchar4 *q0 = colors ? reinterpret_cast<uchar4 *>(c0) + offset : nullptr;
        uchar4 *q1 = colors ? reinterpret_cast<uchar4 *>(c1) + offset : nullptr;
        uchar4 *q2 = colors ? reinterpret_cast<uchar4 *>(c2) + offset : nullptr;
        pmkd::add_mesh_attributes(ctx, reinterpret_cast<const float4 *>(normals),
                                  reinterpret_cast<const uchar4 *>(colors), nv, indices, nidx, model,
                                  reinterpret_cast<float4 *>(n0) + offset, reinterpret_cast<float4 *>(n1) + offset,
                                  reinterpret_cast<float4 *>(n2) + offset, q0, q1, q2, (hipStream_t)stream);
    });
}

int pm_osc_surface(pm_ctx *ctx, float *vertices, float *normals, int detail, float height,
                   float time, float amplitude, float frequency, void *stream) {
    return guarded(ctx, [&] {
        if (detail < 0 || detail >= 46340 || !vertices) throw Error{PM_ERR_ARG, "invalid grid"};
        pmkd::osc_surface(reinterpret_cast<float4 *>(vertices), reinterpret_cast<float4 *>(normals), detail + 1,
                          height, time, amplitude, frequency, (hipStream_t)stream);
    });
}

int pm_deform_radial(pm_ctx *ctx, const float *unit, int nv, float frame, const float *wave,
                     float *out, void *stream) {
    return guarded(ctx, [&] {
        if (nv < 0 || !wave || (nv > 0 && (!unit || !out))) throw Error{PM_ERR_ARG, "invalid arguments"};
        pmkd::radial_wave(reinterpret_cast<const float4 *>(unit), nv, frame, wave, reinterpret_cast<float4 *>(out),
                          (hipStream_t)stream);
    });
}

int pm_create_rays(pm_ctx *ctx, const float *inv_view_proj, const float *cam_pos, int width, int height,
                   unsigned long long jitter_seed, int jitter, float *org, float *dir, void *stream) {
    return guarded(ctx, [&] {
        if (width < 0 || height < 0 || !inv_view_proj || !cam_pos)
            throw Error{PM_ERR_ARG, "invalid camera"};
        const long long n = (long long)width * height;
        if (n > INT32_MAX) throw Error{PM_ERR_ARG, "width * height exceeds 2^31 - 1 rays"};
        if (n > 0 && (!org || !dir)) throw Error{PM_ERR_ARG, "invalid ray arrays"};
        pmkd::create_rays(inv_view_proj, cam_pos, width, height, jitter_seed, jitter != 0,
                          reinterpret_cast<float4 *>(org), reinterpret_cast<float4 *>(dir), (hipStream_t)stream);
    });
}

int pm_photon_emit(pm_ctx *ctx, const float *l0, const float *l1, const float *l2, int num_lights, int n,
                   unsigned long long seed, float offset, float *out_org, float *out_dir, void *stream) {
    return guarded(ctx, [&] {
        if (n < 0) throw Error{PM_ERR_ARG, "invalid photon count"};
        if (n > 0 && (num_lights < 1 || num_lights > PM_MAX_LIGHTS || !l0 || !l1 || !l2 || !out_org || !out_dir))
            throw Error{PM_ERR_ARG, "need 1.." + std::to_string(PM_MAX_LIGHTS) + " light triangles and outputs"};
        pmkd::photon_emit(ctx, reinterpret_cast<const float4 *>(l0), reinterpret_cast<const float4 *>(l1),
                          reinterpret_cast<const float4 *>(l2), num_lights, n, seed, offset,
                          reinterpret_cast<float4 *>(out_org), reinterpret_cast<float4 *>(out_dir),
                          (hipStream_t)stream);
    });
}

int pm_photon_bounce(pm_ctx *ctx, const float *org, const float *dir, const pm_hit *hits, int n,
                     int absorb_from, unsigned long long seed, float offset, float *out_org,
                     float *out_dir, int *out_src, int *out_count, int *n_out, void *stream) {
    return guarded(ctx, [&] {
        if (n < 0 || (n > 0 && (!org || !dir || !hits || !out_org || !out_dir)))
            throw Error{PM_ERR_ARG, "invalid rays"};
        if (!ctx->p0 && ctx->ntris > 0) throw Error{PM_ERR_NO_TREE, "no geometry"};
        if (absorb_from < 0 || absorb_from > ctx->ntris)
            throw Error{PM_ERR_ARG, "absorb_from must be in [0, triangles of the geometry]"};
        // the compaction's scan scratch is shared with a captured batch
        pmkd::sync_graph_state(ctx, (hipStream_t)stream);
        const int c = pmkd::photon_bounce(ctx, reinterpret_cast<const float4 *>(org),
                                          reinterpret_cast<const float4 *>(dir), hits, n, absorb_from,
                                          seed, offset, reinterpret_cast<float4 *>(out_org),
                                          reinterpret_cast<float4 *>(out_dir), out_src, out_count,
                                          (hipStream_t)stream);
        if (n_out) *n_out = c;
    });
}

}  // extern "C"
