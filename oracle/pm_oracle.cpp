/*
 * pm_oracle.cpp -- CPU oracle for the kd-tree build + traversal path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Compiled with
 *   g++ -O3 -std=c++17 -ffp-contract=off -fopenmp
 * so every float operation is an individually rounded IEEE-754 binary32
 * operation in source order.  The product kernels are compiled the same way
 * (-ffp-contract=off, correctly rounded division and square root), which is
 * what makes hit ids and tree arrays comparable bit for bit.
 *
 * Each function cites the reference file:line it restates.  Deliberate
 * departures from the reference (each one also listed in the project notes):
 *  - IEEE 1.0f/det instead of the device __fdividef (IRayTracer.h:125-129);
 *  - box min/max use the IEEE total order (-0 < +0) so that the result does
 *    not depend on evaluation order;
 *  - brute force has no 4096-triangle cap unless brute_cap asks for one
 *    (BruteTracer.cu:50,62);
 *  - the short-stack tracer intersects with the ray direction itself, not
 *    with 1/(1/d) (ShortStack.cu:134,148);
 *  - the upper build stops with an error after max_upper_levels levels or
 *    when a level holds more than max_ref_factor * N + 65536 references,
 *    instead of looping forever (and doubling its references every level)
 *    on >= lower_size coincident triangles;
 *  - fields the reference leaves unwritten (boxes/prim index of empty-space
 *    nodes, w of node boxes, surface area of inner nodes) are zero, and the
 *    parent of a lower SAH child is its parent (LowerTriangleMap.h:557-600
 *    leaves it unset);
 *  - traversal gives up on a ray after more leaf visits than the tree has
 *    nodes or more than 4096 descent steps (only reachable with a corrupt
 *    tree); such a ray reports a miss.
 */
#include "pm_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

/* Utils.h:18: '#define fInfinity 0x7f800000' is an integer literal; every
 * use converts it to the float 2139095040.0f, so that is the sentinel. */
constexpr float F_INF = 2139095040.0f;
constexpr int8_t LEAF = 0, AX_X = 1, AX_Y = 2, AX_Z = 3;   /* KDNode.h:62-65 */
constexpr int MAX_DESCENT = 4096;

thread_local std::string g_err;

struct f3 { float x, y, z; };
struct f4 { float x, y, z, w; };

inline f3 mk3(const f4 &a) { return {a.x, a.y, a.z}; }
inline f3 sub(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
/* cutil_math cross/dot, evaluated left to right */
inline f3 cross(f3 a, f3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline float comp(f3 a, int axis) { return axis == AX_X ? a.x : (axis == AX_Y ? a.y : a.z); }
inline float comp4(const f4 &a, int axis) { return axis == AX_X ? a.x : (axis == AX_Y ? a.y : a.z); }

inline int32_t fbits(float f) { int32_t b; std::memcpy(&b, &f, 4); return b; }
inline float bitsf(int32_t b) { float f; std::memcpy(&f, &b, 4); return f; }
/* order-preserving key: signed compare of keys == IEEE total order */
inline int32_t fkey(float f) { int32_t b = fbits(f); return b >= 0 ? b : (b ^ 0x7fffffff); }
inline float omin(float a, float b) { return fkey(b) < fkey(a) ? b : a; }
inline float omax(float a, float b) { return fkey(b) > fkey(a) ? b : a; }

inline int ctz64(uint64_t x) { return __builtin_ctzll(x); }
inline int popc64(uint64_t x) { return __builtin_popcountll(x); }

/* IRayTracer.h:60-65 */
inline f3 adjust_dir(f3 d) {
    const float e = 0x1p-80f;  /* exp2f(-80.0) */
    d.x = std::fabs(d.x) > e ? d.x : std::copysign(e, d.x);
    d.y = std::fabs(d.y) > e ? d.y : std::copysign(e, d.y);
    d.z = std::fabs(d.z) > e ? d.z : std::copysign(e, d.z);
    return d;
}

/* IRayTracer.h:111-145 (host branch: IEEE division).  Updates (th, tu, tv)
 * and returns true when the triangle is accepted: 0 <= t <= th, u >= 0,
 * v >= 0, u + v <= 1.  Equal t is accepted, so later prims win ties. */
inline bool moller_trumbore(f3 v0, f3 v1, f3 v2, f3 o, f3 d,
                            float &th, float &tu, float &tv) {
    const f3 e1 = sub(v1, v0);
    const f3 e2 = sub(v2, v0);
    const f3 p = cross(d, e2);
    const f3 t = sub(o, v0);
    const f3 q = cross(t, e1);
    const float inv = 1.0f / dot(p, e1);
    const float hx = inv * dot(q, e2);
    if (0.0f <= hx && hx <= th) {
        const float hy = inv * dot(p, t);
        const float hz = inv * dot(q, d);
        if (hy >= 0.0f && hz >= 0.0f && hy + hz <= 1.0f) {
            th = hx; tu = hy; tv = hz;
            return true;
        }
    }
    return false;
}

/* IRayTracer.h:147-168 with invDir = true */
inline bool ray_box(f3 o, f3 inv, f3 bmin, f3 bmax) {
    bmin = {(bmin.x - o.x) * inv.x, (bmin.y - o.y) * inv.y, (bmin.z - o.z) * inv.z};
    bmax = {(bmax.x - o.x) * inv.x, (bmax.y - o.y) * inv.y, (bmax.z - o.z) * inv.z};
    float nearA = std::fmin(bmin.x, bmax.x);
    nearA = std::fmax(nearA, std::fmin(bmin.y, bmax.y));
    nearA = std::fmax(nearA, std::fmin(bmin.z, bmax.z));
    float farA = std::fmax(bmax.x, bmin.x);
    farA = std::fmin(farA, std::fmax(bmin.y, bmax.y));
    farA = std::fmin(farA, std::fmax(bmin.z, bmax.z));
    return nearA <= farA;
}

/* Akenine-Moller triangle/box step 3 (the 9 edge x axis separating axes),
 * IntersectionTests.h:20-160,378-395.  The three a0X/a1X/a2X groups are
 * the same 2-D test on the (y,z), (z,x) and (x,y) components; GAMMA and
 * EPSILON widen each interval. */
constexpr float SAT_GAMMA = 0.00001f, SAT_EPS = 1.0001f;

inline bool sat_edge(float au, float aw, float bu, float bw, float cu, float cw, float hu, float hw) {
    const float p0 = aw * bu - au * bw;
    const float p1 = (bu - au) * cw - (bw - aw) * cu;
    float r = hw * std::fabs(bu - au) + hu * std::fabs(bw - aw) + SAT_GAMMA;
    r *= SAT_EPS;
    return !((p0 > r && p1 > r) || (p0 < -r && p1 < -r));
}

inline bool sat_plane(float u0, float w0, float u1, float w1, float u2, float w2, float hu, float hw) {
    bool ret = true;
    ret &= sat_edge(u0, w0, u1, w1, u2, w2, hu, hw);
    ret &= sat_edge(u1, w1, u2, w2, u0, w0, hu, hw);
    ret &= sat_edge(u2, w2, u0, w0, u1, w1, hu, hw);
    return ret;
}

inline bool tri_box_step3(f3 v0, f3 v1, f3 v2, f3 bmin, f3 bmax) {
    const f3 h = {(bmax.x - bmin.x) * 0.5f, (bmax.y - bmin.y) * 0.5f, (bmax.z - bmin.z) * 0.5f};
    const f3 c = {bmin.x + h.x, bmin.y + h.y, bmin.z + h.z};
    v0 = sub(v0, c); v1 = sub(v1, c); v2 = sub(v2, c);
    bool ret = true;
    ret &= sat_plane(v0.y, v0.z, v1.y, v1.z, v2.y, v2.z, h.y, h.z);   /* a00..a02 */
    ret &= sat_plane(v0.z, v0.x, v1.z, v1.x, v2.z, v2.x, h.z, h.x);   /* a10..a12 */
    ret &= sat_plane(v0.x, v0.y, v1.x, v1.y, v2.x, v2.y, h.x, h.y);   /* a20..a22 */
    return ret;
}

/* WoopTransformationMatrix, IntersectionTests.h:437-467 */
inline void woop_matrix(f3 a, f3 b, f3 c, f4 &m0, f4 &m1, f4 &m2) {
    const f3 i0 = sub(a, c), i1 = sub(b, c), i2 = sub(cross(i0, i1), c);
    const float det = i0.x * (i1.y * i2.z - i1.z * i2.y) + i1.x * (i2.y * i0.z - i2.z * i0.y) +
                      i2.x * (i0.y * i1.z - i0.z * i1.y);
    const float inv = 1.0f / det;
    const f3 t0 = {(i1.y * i2.z - i1.z * i2.y) * inv, (i2.x * i1.z - i2.z * i1.x) * inv,
                   (i1.x * i2.y - i1.y * i2.x) * inv};
    const f3 t1 = {(i2.y * i0.z - i2.z * i0.y) * inv, (i0.x * i2.z - i0.z * i2.x) * inv,
                   (i2.x * i0.y - i2.y * i0.x) * inv};
    const f3 t2 = {(i0.y * i1.z - i0.z * i1.y) * inv, (i1.x * i0.z - i1.z * i0.x) * inv,
                   (i0.x * i1.y - i0.y * i1.x) * inv};
    m0 = {t0.x, t0.y, t0.z, dot(t0, c)};
    m1 = {t1.x, t1.y, t1.z, dot(t1, c)};
    m2 = {t2.x, t2.y, t2.z, dot(t2, c)};
}

/* IRayTracer::Woop with WoopLambda / WoopUV, IRayTracer.h:78-109 */
inline bool woop_test(const f4 &m0, const f4 &m1, const f4 &m2, f3 o, f3 d, float &th, float &tu, float &tv) {
    const f3 n = mk3(m2);
    const float lambda = -(dot(n, o) - m2.w) / dot(n, d);
    if (0.0f <= lambda && lambda <= th) {
        const f3 a = mk3(m0), b = mk3(m1);
        const float u = lambda * dot(a, d) + dot(a, o) - m0.w;
        const float v = lambda * dot(b, d) + dot(b, o) - m1.w;
        if (u >= -0.0f && v >= -0.0f && u + v <= 1.0f) {
            th = lambda; tu = u; tv = v;
            return true;
        }
    }
    return false;
}

/* LowerTriangleMap.h:32-36: 0.5f * length(cross(v1-v0, v2-v0)) */
inline float tri_area(f3 v0, f3 v1, f3 v2) {
    const f3 c = cross(sub(v1, v0), sub(v2, v0));
    return 0.5f * std::sqrt(dot(c, c));
}

struct Tree {
    std::vector<int8_t> info;
    std::vector<float> split;
    std::vector<f4> bmin, bmax;
    std::vector<int> prim_index;
    std::vector<int64_t> bitmap;   /* count for inner upper nodes, mask for leaves */
    std::vector<int> children;     /* 2 per node */
    std::vector<float> area;
    std::vector<int> parent;
    std::vector<int> prim_indices;
    std::vector<int> leaf_ids;
    std::vector<long long> level_refs;
    std::vector<long long> lower_level_nodes;
    int upper_levels = 0, lower_levels = 0;
    int num_roots = 1;             /* roots are nodes 0..num_roots-1 */

    int size() const { return (int)info.size(); }
    void resize(int n) {
        info.resize(n, 0); split.resize(n, 0.0f);
        bmin.resize(n, f4{0, 0, 0, 0}); bmax.resize(n, f4{0, 0, 0, 0});
        prim_index.resize(n, 0); bitmap.resize(n, 0); children.resize(2 * (size_t)n, 0);
        area.resize(n, 0.0f); parent.resize(n, 0);
    }
};

struct Params {
    int lower_algorithm, empty_space_splitting, split_method, propagate_aabbs;
    float empty_space_threshold, traversal_cost, sah_empty_bonus;
    int lower_size, max_upper_levels, sah_leaf_max, max_ref_factor;
};

template <class T>
void exclusive_scan(const std::vector<T> &in, std::vector<T> &out) {
    out.resize(in.size());
    T acc = 0;
    for (size_t i = 0; i < in.size(); ++i) { out[i] = acc; acc += in[i]; }
}

/* ------------------------------------------------------------------ */
/* Upper (large node) stage: TriangleMapUpperCreator.cu:110-759        */
/* ------------------------------------------------------------------ */
int build_upper(const f4 *P0, const f4 *P1, const f4 *P2, int N, const int *off, int M,
                const Params &prm, Tree &T, std::vector<f4> &prim_min, std::vector<f4> &prim_max) {
    const int LS = prm.lower_size;
    const int max_levels = prm.max_upper_levels > 0 ? prm.max_upper_levels : 128;

    /* root: TriangleMapUpperCreator.cu:128-137.  A forest of M meshes has
     * one root per mesh, root m (node m) over triangles [off[m], off[m+1])
     * and its own parent; the level loop below treats them as the first
     * level's M active nodes. */
    T.resize(M);
    T.num_roots = M;
    for (int m = 0; m < M; ++m) {
        T.prim_index[m] = off[m]; T.bitmap[m] = off[m + 1] - off[m]; T.parent[m] = m;
    }

    /* CalcPrimitiveAabb, Kernels/TriangleUpper.h:10-28 */
    std::vector<f4> rmin((size_t)N), rmax((size_t)N);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; ++i) {
        f3 a = mk3(P0[i]), b = mk3(P1[i]), c = mk3(P2[i]);
        f3 mn = a, mx = a;
        mn = {omin(mn.x, b.x), omin(mn.y, b.y), omin(mn.z, b.z)};
        mx = {omax(mx.x, b.x), omax(mx.y, b.y), omax(mx.z, b.z)};
        mn = {omin(mn.x, c.x), omin(mn.y, c.y), omin(mn.z, c.z)};
        mx = {omax(mx.x, c.x), omax(mx.y, c.y), omax(mx.z, c.z)};
        rmin[i] = {mn.x, mn.y, mn.z, bitsf(i)};   /* __int_as_float(id) */
        rmax[i] = {mx.x, mx.y, mx.z, 0.0f};
    }

    int active_index = 0, A = M, level = 0;
    std::vector<f4> prop_min, prop_max, nmin, nmax;
    std::vector<int> side, addr, leaf_side, leaf_addr, lf, lna, owner;
    while (A > 0) {
        if (level >= max_levels) {
            g_err = "upper build did not finish within " + std::to_string(max_levels) +
                    " levels (>= lower_size coincident triangles?)";
            return 3;
        }
        const int R = (int)rmin.size();
        const long long max_refs = (long long)(prm.max_ref_factor > 0 ? prm.max_ref_factor : 16) * N + 65536;
        if (R > max_refs) {
            g_err = "upper build exceeded " + std::to_string(max_refs) +
                    " references (>= lower_size coincident triangles?)";
            return 3;
        }
        T.level_refs.push_back(R);

        /* ReduceAabb, TriangleMapUpperCreator.cu:241-318 */
        prop_min.assign(A, f4{0, 0, 0, 0}); prop_max.assign(A, f4{0, 0, 0, 0});
        if (prm.empty_space_splitting)
            for (int i = 0; i < A; ++i) {
                prop_min[i] = T.bmin[active_index + i];
                prop_max[i] = T.bmax[active_index + i];
            }
        owner.assign(R, 0);
#pragma omp parallel for schedule(dynamic, 64)
        for (int i = 0; i < A; ++i) {
            const int node = active_index + i;
            const int pi = T.prim_index[node];
            const int cnt = (int)T.bitmap[node];
            f3 mn = {F_INF, F_INF, F_INF}, mx = {-F_INF, -F_INF, -F_INF}; /* AabbMemset */
            for (int k = pi; k < pi + cnt; ++k) {
                owner[k] = i;
                mn = {omin(mn.x, rmin[k].x), omin(mn.y, rmin[k].y), omin(mn.z, rmin[k].z)};
                mx = {omax(mx.x, rmax[k].x), omax(mx.y, rmax[k].y), omax(mx.z, rmax[k].z)};
            }
            T.bmin[node] = {mn.x, mn.y, mn.z, 0.0f};
            T.bmax[node] = {mx.x, mx.y, mx.z, 0.0f};
        }

        /* CreateEmptySplits, TriangleMapUpperCreator.cu:320-368;
         * CalcEmptySpaceSplits / EmptySpaceSplitting2, Kernels/EmptySpaceSplitting.h:10-184 */
        if (prm.empty_space_splitting && active_index != 0) {
            const float thr = prm.empty_space_threshold;
            std::vector<int> planes(A, 0), req(A, 0), eaddr;
            bool any = false;
            for (int i = 0; i < A; ++i) {
                const f3 pMin = mk3(prop_min[i]), pMax = mk3(prop_max[i]);
                const f3 rMin = mk3(T.bmin[active_index + i]), rMax = mk3(T.bmax[active_index + i]);
                int pl = 0;
                if (rMin.x - pMin.x > thr * (pMax.x - pMin.x)) pl |= 1;
                if (rMin.y - pMin.y > thr * (pMax.y - pMin.y)) pl |= 2;
                if (rMin.z - pMin.z > thr * (pMax.z - pMin.z)) pl |= 4;
                if (pMax.x - rMax.x > thr * (pMax.x - pMin.x)) pl |= 8;
                if (pMax.y - rMax.y > thr * (pMax.y - pMin.y)) pl |= 16;
                if (pMax.z - rMax.z > thr * (pMax.z - pMin.z)) pl |= 32;
                planes[i] = pl; req[i] = 2 * popc64((uint64_t)pl);
                any |= pl != 0;
            }
            if (any) {
                exclusive_scan(req, eaddr);
                const int E = eaddr[A - 1] + req[A - 1];
                const int base = T.size();
                T.resize(base + E);
                for (int i = 0; i < A; ++i) {
                    int pl = planes[i];
                    if (!pl) continue;
                    const int leafID = active_index + i;
                    const int parentID = T.parent[leafID];
                    const int start = active_index + A + eaddr[i];
                    if (T.children[2 * parentID] == leafID) T.children[2 * parentID] = start;
                    else T.children[2 * parentID + 1] = start;
                    const f3 minP = mk3(T.bmin[leafID]), maxP = mk3(T.bmax[leafID]);
                    int pid = start;
                    const int8_t lo_axis[3] = {AX_X, AX_Y, AX_Z};
                    const float lo_pos[3] = {minP.x, minP.y, minP.z};
                    for (int a = 0; a < 3; ++a) {
                        if (!(pl & (1 << a))) continue;
                        pl -= 1 << a;
                        T.info[pid] = lo_axis[a]; T.split[pid] = lo_pos[a];
                        const int left = pid + 1, right = pl ? left + 1 : leafID;
                        T.children[2 * pid] = left; T.children[2 * pid + 1] = right;
                        T.info[left] = LEAF; T.bitmap[left] = 0;
                        pid = right;
                    }
                    /* max planes; the max-z case emits an X split at max.x
                     * exactly as EmptySpaceSplitting.h:172-175 does */
                    const int8_t hi_axis[3] = {AX_X, AX_Y, AX_X};
                    const float hi_pos[3] = {maxP.x, maxP.y, maxP.x};
                    for (int a = 0; a < 3; ++a) {
                        if (!(pl & (8 << a))) continue;
                        pl -= 8 << a;
                        T.info[pid] = hi_axis[a]; T.split[pid] = hi_pos[a];
                        const int right = pid + 1, left = pl ? right + 1 : leafID;
                        T.children[2 * pid] = left; T.children[2 * pid + 1] = right;
                        T.info[right] = LEAF; T.bitmap[right] = 0;
                        pid = left;
                    }
                }
            }
        }

        /* CalcUpperNodeSplitInfo, Kernels/ReduceSegments.h:466-484 */
        for (int i = 0; i < A; ++i) {
            const int node = active_index + i;
            const f3 mn = mk3(T.bmin[node]), mx = mk3(T.bmax[node]);
            const f3 sz = sub(mx, mn);
            const f3 med = {sz.x * 0.5f + mn.x, sz.y * 0.5f + mn.y, sz.z * 0.5f + mn.z};
            const bool yAboveX = sz.x < sz.y;
            const float mxs = yAboveX ? sz.y : sz.x;
            const float spl = yAboveX ? med.y : med.x;
            const int8_t ax = yAboveX ? AX_Y : AX_X;
            const bool zHigher = mxs < sz.z;
            T.split[node] = zHigher ? med.z : spl;
            T.info[node] = zHigher ? AX_Z : ax;
        }

        /* CreateChildren, TriangleMapUpperCreator.cu:467-759 */
        const int child_start = T.size();
        T.resize(child_start + 2 * A);

        /* SetSplitSide, Kernels/TriangleUpperChildren.h:10-62 */
        side.assign(2 * (size_t)R + 1, 0);
#pragma omp parallel for schedule(static)
        for (int k = 0; k < R; ++k) {
            const int node = active_index + owner[k];
            const int ax = T.info[node];
            const float s = T.split[node];
            bool l = comp4(rmin[k], ax) <= s, r = s < comp4(rmax[k], ax);
            if (prm.split_method == 1 && l && r) {
                /* SetDivideSide<false>, TriangleUpperChildren.h:64-142: a
                 * straddling triangle goes to the child boxes it meets */
                const int t = fbits(rmin[k].w);
                const f3 v0 = mk3(P0[t]), v1 = mk3(P1[t]), v2 = mk3(P2[t]);
                const f3 nmn = mk3(T.bmin[node]), nmx = mk3(T.bmax[node]);
                const f3 lmax = {ax == AX_X ? s : nmx.x, ax == AX_Y ? s : nmx.y, ax == AX_Z ? s : nmx.z};
                const f3 rmn = {ax == AX_X ? s : nmn.x, ax == AX_Y ? s : nmn.y, ax == AX_Z ? s : nmn.z};
                l = tri_box_step3(v0, v1, v2, nmn, lmax);
                r = tri_box_step3(v0, v1, v2, rmn, nmx);
            }
            side[k] = l;
            side[(size_t)R + k] = r;
        }
        exclusive_scan(side, addr);
        const int newTri = addr[2 * (size_t)R];
        if (newTri < R) {   /* TriangleMapUpperCreator.cu:606-607 */
            g_err = "new triangle count below old";
            return 4;
        }
        /* CalcNodeChildSize, TriangleUpperChildren.h:144-162 */
        std::vector<int> Lsz(A), Rsz(A);
        bool created_leafs = false;
        for (int i = 0; i < A; ++i) {
            const int node = active_index + i;
            const int pi = T.prim_index[node], cnt = (int)T.bitmap[node];
            Lsz[i] = addr[pi + cnt] - addr[pi];
            Rsz[i] = addr[(size_t)pi + cnt + R] - addr[(size_t)pi + R];
            if (Lsz[i] < LS || Rsz[i] < LS) created_leafs = true;
        }
        int children_created;
        if (created_leafs) {
            /* SetPrimitiveLeafSide, TriangleUpperChildren.h:164-186 */
            leaf_side.assign(2 * (size_t)R + 1, 0);
#pragma omp parallel for schedule(static)
            for (int k = 0; k < R; ++k) {
                leaf_side[k] = (Lsz[owner[k]] < LS) * side[k];
                leaf_side[(size_t)R + k] = (Rsz[owner[k]] < LS) * side[(size_t)R + k];
            }
            exclusive_scan(leaf_side, leaf_addr);
            const int leafTri = leaf_addr[2 * (size_t)R];
            const int nextR = newTri - leafTri;
            const int ulp = (int)prim_min.size();
            prim_min.resize((size_t)ulp + leafTri); prim_max.resize((size_t)ulp + leafTri);
            nmin.assign(nextR, f4{0, 0, 0, 0}); nmax.assign(nextR, f4{0, 0, 0, 0});
            /* SplitTriangles (leaf version), TriangleUpperChildren.h:329-383 */
#pragma omp parallel for schedule(static)
            for (int k = 0; k < R; ++k) {
                const int node = active_index + owner[k];
                const int ax = T.info[node];
                const float s = T.split[node];
                const f4 mn = rmin[k], mx = rmax[k];
                if (side[k]) {
                    const f4 cmx = {ax == AX_X ? omin(s, mx.x) : mx.x, ax == AX_Y ? omin(s, mx.y) : mx.y,
                                    ax == AX_Z ? omin(s, mx.z) : mx.z, mx.w};
                    if (leaf_side[k]) { prim_min[ulp + leaf_addr[k]] = mn; prim_max[ulp + leaf_addr[k]] = cmx; }
                    else { const int a = addr[k] - leaf_addr[k]; nmin[a] = mn; nmax[a] = cmx; }
                }
                const size_t kr = (size_t)R + k;
                if (side[kr]) {
                    const f4 cmn = {ax == AX_X ? omax(s, mn.x) : mn.x, ax == AX_Y ? omax(s, mn.y) : mn.y,
                                    ax == AX_Z ? omax(s, mn.z) : mn.z, mn.w};
                    if (leaf_side[kr]) { prim_min[ulp + leaf_addr[kr]] = cmn; prim_max[ulp + leaf_addr[kr]] = mx; }
                    else { const int a = addr[kr] - leaf_addr[kr]; nmin[a] = cmn; nmax[a] = mx; }
                }
            }
            /* MarkNodeLeafs + scan, TriangleUpperChildren.h:190-201, .cu:662-671 */
            lf.assign(2 * (size_t)A + 1, 0);
            for (int i = 0; i < A; ++i) { lf[i] = Lsz[i] < LS; lf[A + i] = Rsz[i] < LS; }
            exclusive_scan(lf, lna);
            const int leaf_nodes = lna[2 * A];
            /* CreateUpperChildren (leaf version), TriangleUpperChildren.h:237-281 */
            for (int i = 0; i < A; ++i) {
                const int par = active_index + i;
                const int pi = T.prim_index[par];
                bool isLeaf = Lsz[i] < LS;
                const int leftID = child_start + (isLeaf ? lna[i] : i - lna[i] + leaf_nodes);
                T.prim_index[leftID] = isLeaf ? leaf_addr[pi] + ulp : addr[pi] - leaf_addr[pi];
                T.bitmap[leftID] = Lsz[i]; T.parent[leftID] = par;
                const int j = A + i;
                isLeaf = Rsz[i] < LS;
                const int rightID = child_start + (isLeaf ? lna[j] : j - lna[j] + leaf_nodes);
                const size_t pr = (size_t)pi + R;
                T.prim_index[rightID] = isLeaf ? leaf_addr[pr] + ulp : addr[pr] - leaf_addr[pr];
                T.bitmap[rightID] = Rsz[i]; T.parent[rightID] = par;
                T.children[2 * par] = leftID; T.children[2 * par + 1] = rightID;
            }
            children_created = 2 * A - leaf_nodes;
            /* MarkLeafNodes, TriangleUpperChildren.h:385-395 */
            for (int i = 0; i < leaf_nodes; ++i) T.leaf_ids.push_back(child_start + i);
        } else {
            nmin.assign(newTri, f4{0, 0, 0, 0}); nmax.assign(newTri, f4{0, 0, 0, 0});
            /* CreateUpperChildren (plain), TriangleUpperChildren.h:203-234 */
            for (int i = 0; i < A; ++i) {
                const int par = active_index + i;
                const int pi = T.prim_index[par];
                const int leftID = child_start + i, rightID = leftID + A;
                T.prim_index[leftID] = addr[pi];
                T.prim_index[rightID] = addr[(size_t)pi + R];
                T.bitmap[leftID] = Lsz[i]; T.bitmap[rightID] = Rsz[i];
                T.children[2 * par] = leftID; T.children[2 * par + 1] = rightID;
                T.parent[leftID] = T.parent[rightID] = par;
            }
            /* SplitTriangles (plain), TriangleUpperChildren.h:283-327 */
#pragma omp parallel for schedule(static)
            for (int k = 0; k < R; ++k) {
                const int node = active_index + owner[k];
                const int ax = T.info[node];
                const float s = T.split[node];
                const f4 mn = rmin[k], mx = rmax[k];
                if (side[k]) {
                    const int a = addr[k];
                    nmin[a] = mn;
                    nmax[a] = {ax == AX_X ? omin(s, mx.x) : mx.x, ax == AX_Y ? omin(s, mx.y) : mx.y,
                               ax == AX_Z ? omin(s, mx.z) : mx.z, mx.w};
                }
                const size_t kr = (size_t)R + k;
                if (side[kr]) {
                    const int a = addr[kr];
                    nmin[a] = {ax == AX_X ? omax(s, mn.x) : mn.x, ax == AX_Y ? omax(s, mn.y) : mn.y,
                               ax == AX_Z ? omax(s, mn.z) : mn.z, mn.w};
                    nmax[a] = mx;
                }
            }
            children_created = 2 * A;
        }

        /* PropagateAabbToChildren<false>, Kernels/EmptySpaceSplitting.h:299-331 */
        if (prm.empty_space_splitting) {
            for (int i = 0; i < A; ++i) {
                const int par = active_index + i;
                const int ax = T.info[par] & 3;
                const float s = T.split[par];
                const int l = T.children[2 * par], r = T.children[2 * par + 1];
                const f4 mn = T.bmin[par], mx = T.bmax[par];
                T.bmin[l] = mn;
                T.bmin[r] = {ax == AX_X ? s : mn.x, ax == AX_Y ? s : mn.y, ax == AX_Z ? s : mn.z, 0.0f};
                T.bmax[l] = {ax == AX_X ? s : mx.x, ax == AX_Y ? s : mx.y, ax == AX_Z ? s : mx.z, 0.0f};
                T.bmax[r] = mx;
            }
        }
        rmin.swap(nmin); rmax.swap(nmax);
        active_index = T.size() - children_created;
        A = children_created;
        ++level;
    }
    T.upper_levels = level;

    /* ExtractIndexFromAabb, Kernels/TriangleUpper.h:43-52 */
    T.prim_indices.resize(prim_min.size());
    for (size_t k = 0; k < prim_min.size(); ++k) T.prim_indices[k] = fbits(prim_min[k].w);

    /* PropagateParentAabb<true>, EmptySpaceSplitting.h:333-368 (.cu:173-182) */
    if (!prm.empty_space_splitting && prm.propagate_aabbs) {
        for (int leaf : T.leaf_ids) {
            const int par = T.parent[leaf];
            const f3 pmn = mk3(T.bmin[par]), pmx = mk3(T.bmax[par]);
            const int ax = T.info[par] & 3;
            const float s = T.split[par];
            if (T.children[2 * par] == leaf) {
                T.bmin[leaf] = {pmn.x, pmn.y, pmn.z, 0.0f};
                T.bmax[leaf] = {ax == AX_X ? s : pmx.x, ax == AX_Y ? s : pmx.y, ax == AX_Z ? s : pmx.z, 0.0f};
            } else {
                T.bmin[leaf] = {ax == AX_X ? s : pmn.x, ax == AX_Y ? s : pmn.y, ax == AX_Z ? s : pmn.z, 0.0f};
                T.bmax[leaf] = {pmx.x, pmx.y, pmx.z, 0.0f};
            }
        }
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* Lower SAH decision for one node: CalcSAH + CalcAreaForSets,         */
/* Kernels/LowerTriangleMap.h:412-555                                  */
/* ------------------------------------------------------------------ */
struct SahResult {
    bool split; int axis; float pos; uint64_t left, right;
    float left_area, right_area, cost;
    int split_slot; bool high;
};

/* slot i: box (mins[i], maxs[i]) and area areas[i]; only bits of bmp used */
SahResult sah_node(const f3 *mins, const f3 *maxs, const float *areas, uint64_t bmp,
                   float node_area, float Ct, int leaf_max, float bonus) {
    SahResult r{};
    r.split = false; r.axis = LEAF; r.cost = F_INF; r.split_slot = -1; r.high = false;
    const int n = popc64(bmp);
    const bool forced = leaf_max > 0 && n > leaf_max;
    float opt = F_INF;
    bool have = false;
    auto consider = [&](uint64_t L, uint64_t R, int axis, int slot, bool high) {
        float aL = 0.0f, aR = 0.0f;
        for (uint64_t b = bmp; b; b &= b - 1) {       /* ascending bit order */
            const int i = ctz64(b);
            aL += (L >> i) & 1 ? areas[i] : 0.0f;
            aR += (R >> i) & 1 ? areas[i] : 0.0f;
        }
        float cost = (float)popc64(L) * aL + (float)popc64(R) * aR;
        if (bonus != 1.0f && (L == 0 || R == 0)) cost = cost * bonus;
        const bool progress = L != bmp && R != bmp;
        if (forced && !progress) return;
        if (cost < opt) {
            opt = cost; have = true;
            r.left = L; r.right = R; r.left_area = aL; r.right_area = aR;
            r.axis = axis; r.split_slot = slot; r.high = high;
        }
    };
    for (uint64_t tb = bmp; tb; tb &= tb - 1) {
        const int i = ctz64(tb);
        for (int a = AX_X; a <= AX_Z; ++a) {
            const float lo = comp(mins[i], a), hi = comp(maxs[i], a);
            uint64_t lL = 0, lR = 0, hL = 0, hR = 0;
            for (uint64_t b = bmp; b; b &= b - 1) {     /* CreateSplittingPlanes, LowerTriangleMap.h:124-146 */
                const int j = ctz64(b);
                const float mnj = comp(mins[j], a), mxj = comp(maxs[j], a);
                const uint64_t bit = (uint64_t)1 << j;
                if (mnj <= lo) lL |= bit;
                if (lo < mxj) lR |= bit;
                if (mnj <= hi) hL |= bit;
                if (hi < mxj) hR |= bit;
            }
            consider(lL, lR, a, i, false);
            consider(hL, hR, a, i, true);
        }
    }
    r.cost = opt;
    if (!have) return r;
    const bool progress = r.left != bmp && r.right != bmp;
    const float leaf_cost = ((float)n - Ct) * node_area;
    r.split = progress && (forced || opt < leaf_cost);
    if (r.split) r.pos = r.high ? comp(maxs[r.split_slot], r.axis) : comp(mins[r.split_slot], r.axis);
    return r;
}

/* The balanced ("SSAH") lower creator, TriangleMapBalancedCreator.cu:62-244
 * with CreateSplittingPlanes / CalcRelationForSets / CalcSplit /
 * CreateChildren (LowerTriangleMap.h:84-155,266-410): the candidate planes
 * are SAH's (6 per prim, prim ascending, X/Y/Z, low then high), the cost of
 * a candidate is small^2 + large^2 of its two child counts, the first
 * minimum below lower_size^2 wins, and the node splits iff
 * (C_t + cost) / n < n. */
struct BalancedResult { bool split; int axis; float pos; uint64_t left, right; int cost; };

/* the decision for one node; mins/maxs indexed by bit (slot - prim index) */
BalancedResult balanced_node(const f3 *mins, const f3 *maxs, uint64_t bmp, float Ct, int lower_size) {
    int best = lower_size * lower_size, bax = 0, bslot = 0;
    bool bhigh = false;
    uint64_t bl = 0, br = 0;
    for (uint64_t b = bmp; b; b &= b - 1) {
        const int i = ctz64(b);
        for (int ax = AX_X; ax <= AX_Z; ++ax) {
            const float lo = comp(mins[i], ax), hi = comp(maxs[i], ax);
            uint64_t ll = 0, lr = 0, hl = 0, hr = 0;
            for (uint64_t c = bmp; c; c &= c - 1) {
                const int j = ctz64(c);
                const float mnj = comp(mins[j], ax), mxj = comp(maxs[j], ax);
                const uint64_t bit = (uint64_t)1 << j;
                if (mnj <= lo) ll |= bit;
                if (lo < mxj) lr |= bit;
                if (mnj <= hi) hl |= bit;
                if (hi < mxj) hr |= bit;
            }
            for (int h = 0; h < 2; ++h) {
                const uint64_t l = h ? hl : ll, r = h ? hr : lr;
                const int small = std::min(popc64(l), popc64(r)), large = std::max(popc64(l), popc64(r));
                const int area = small * small + large * large;
                if (area < best) { best = area; bl = l; br = r; bax = ax; bslot = i; bhigh = h; }
            }
        }
    }
    BalancedResult res{};
    const int n = popc64(bmp);
    const float ssah = (Ct + (float)best) / (float)n;
    res.split = n > 0 && ssah < (float)n;
    res.cost = best;
    if (res.split) {
        res.axis = bax; res.left = bl; res.right = br;
        res.pos = bhigh ? comp(maxs[bslot], bax) : comp(mins[bslot], bax);
    }
    return res;
}

int build_lower_balanced(const Params &prm, Tree &T, const std::vector<f4> &prim_min,
                         const std::vector<f4> &prim_max) {
    const int LS = prm.lower_size;
    for (int leaf : T.leaf_ids) {   /* PreprocessLeafNodes, LowerTriangleMap.h:40-54 */
        const int64_t cnt = T.bitmap[leaf];
        T.bitmap[leaf] = cnt >= 64 ? -1 : (int64_t)(((uint64_t)1 << cnt) - 1);
    }
    std::vector<int> list(T.leaf_ids);
    int level = 0;
    while (!list.empty()) {
        const int A = (int)list.size();
        T.lower_level_nodes.push_back(A);
        std::vector<int> flags(A, 0), axis(A, 0);
        std::vector<uint64_t> Ls(A), Rs(A);
        std::vector<float> pos(A);
#pragma omp parallel for schedule(dynamic, 16)
        for (int id = 0; id < A; ++id) {
            const int par = list[id];
            const int pi = T.prim_index[par];
            const uint64_t bmp = (uint64_t)T.bitmap[par];
            f3 mins[64], maxs[64];
            for (uint64_t b = bmp; b; b &= b - 1) {
                const int i = ctz64(b);
                mins[i] = mk3(prim_min[pi + i]); maxs[i] = mk3(prim_max[pi + i]);
            }
            const BalancedResult r = balanced_node(mins, maxs, bmp, prm.traversal_cost, LS);
            flags[id] = r.split;
            if (r.split) { axis[id] = r.axis; Ls[id] = r.left; Rs[id] = r.right; pos[id] = r.pos; }
        }
        std::vector<int> off;
        for (int id = 0; id < A; ++id) {
            const int par = list[id];
            T.info[par] = flags[id] ? (int8_t)axis[id] : LEAF;
            if (flags[id]) T.split[par] = pos[id];
        }
        exclusive_scan(flags, off);
        const int splits = A ? off[A - 1] + flags[A - 1] : 0;
        const int base = T.size();
        T.resize(base + 2 * splits);
        for (int id = 0; id < A; ++id) {   /* CreateChildren, LowerTriangleMap.h:371-410 */
            if (!flags[id]) continue;
            const int par = list[id];
            const int l = base + off[id], r = l + splits;
            T.prim_index[l] = T.prim_index[par]; T.bitmap[l] = (int64_t)Ls[id];
            T.prim_index[r] = T.prim_index[par]; T.bitmap[r] = (int64_t)Rs[id];
            T.parent[l] = T.parent[r] = par;
            T.children[2 * par] = l; T.children[2 * par + 1] = r;
            if (prm.propagate_aabbs) {   /* PropagateAabbToChildren, LowerTriangleMap.h:602-636 */
                const int ax = axis[id];
                const float s = pos[id];
                const f4 mn = T.bmin[par], mx = T.bmax[par];
                T.bmin[l] = mn;
                T.bmin[r] = {ax == AX_X ? s : mn.x, ax == AX_Y ? s : mn.y, ax == AX_Z ? s : mn.z, 0.0f};
                T.bmax[l] = {ax == AX_X ? s : mx.x, ax == AX_Y ? s : mx.y, ax == AX_Z ? s : mx.z, 0.0f};
                T.bmax[r] = mx;
            }
        }
        list.clear();
        for (int i = 0; i < 2 * splits; ++i) list.push_back(base + i);
        ++level;
    }
    T.lower_levels = level;
    return 0;
}

int build_lower(const f4 *P0, const f4 *P1, const f4 *P2, const Params &prm, Tree &T,
                const std::vector<f4> &prim_min, const std::vector<f4> &prim_max) {
    const int LS = prm.lower_size;
    if (prm.lower_algorithm == 0) {
        /* TriangleMapBitmapCreator.cu:32-49 */
        for (int leaf : T.leaf_ids) {
            T.info[leaf] = LEAF;
            const int64_t cnt = T.bitmap[leaf];
            T.bitmap[leaf] = cnt >= 64 ? -1 : (int64_t)(((uint64_t)1 << cnt) - 1);
        }
        return 0;
    }
    if (prm.lower_algorithm == 1) return build_lower_balanced(prm, T, prim_min, prim_max);
    if (prm.lower_algorithm != 2) { g_err = "lower algorithm not supported"; return 5; }
    (void)LS;
    const size_t RL = prim_min.size();
    /* CalcSurfaceArea, LowerTriangleMap.h:24-38 */
    std::vector<float> areas(RL);
#pragma omp parallel for schedule(static)
    for (long long k = 0; k < (long long)RL; ++k) {
        const int t = T.prim_indices[k];
        areas[k] = tri_area(mk3(P0[t]), mk3(P1[t]), mk3(P2[t]));
    }
    /* PreprocesLowerNodes, LowerTriangleMap.h:56-81 */
    for (int leaf : T.leaf_ids) {
        const int pi = T.prim_index[leaf];
        const int cnt = (int)T.bitmap[leaf];
        float a = 0.0f; uint64_t bmp = 0;
        for (int i = 0; i < cnt; ++i) {
            const float ai = areas[pi + i];
            bmp += ai > 0.0f ? ((uint64_t)1 << i) : 0;
            a += ai;
        }
        T.area[leaf] = a;
        T.bitmap[leaf] = (int64_t)bmp;
    }
    /* level loop, TriangleMapSAHCreator.cu:53-86 and ProcessLowerNodes :139-246 */
    std::vector<int> list(T.leaf_ids);
    int level = 0;
    while (!list.empty()) {
        const int A = (int)list.size();
        T.lower_level_nodes.push_back(A);
        std::vector<SahResult> res(A);
#pragma omp parallel for schedule(dynamic, 16)
        for (int id = 0; id < A; ++id) {
            const int par = list[id];
            const int pi = T.prim_index[par];
            const uint64_t bmp = (uint64_t)T.bitmap[par];
            f3 mins[64], maxs[64]; float ar[64];
            for (uint64_t b = bmp; b; b &= b - 1) {
                const int i = ctz64(b);
                mins[i] = mk3(prim_min[pi + i]); maxs[i] = mk3(prim_max[pi + i]); ar[i] = areas[pi + i];
            }
            const float nodeA = T.area[par];
            /* exact shortcut: costs are >= 0, so (n - Ct) * A <= 0 can never be beaten */
            if (prm.sah_leaf_max <= 0 || popc64(bmp) <= prm.sah_leaf_max) {
                if (!(((float)popc64(bmp) - prm.traversal_cost) * nodeA > 0.0f)) {
                    res[id] = SahResult{}; res[id].split = false;
                    continue;
                }
            }
            res[id] = sah_node(mins, maxs, ar, bmp, nodeA, prm.traversal_cost,
                               prm.sah_leaf_max, prm.sah_empty_bonus);
        }
        std::vector<int> flags(A), off;
        for (int id = 0; id < A; ++id) {
            const int par = list[id];
            T.info[par] = res[id].split ? (int8_t)res[id].axis : LEAF;
            if (res[id].split) T.split[par] = res[id].pos;
            flags[id] = res[id].split;
        }
        exclusive_scan(flags, off);
        const int splits = A ? off[A - 1] + flags[A - 1] : 0;
        const int base = T.size();
        T.resize(base + 2 * splits);
        /* CreateLowerSAHChildren, LowerTriangleMap.h:557-600 */
        for (int id = 0; id < A; ++id) {
            if (!flags[id]) continue;
            const int par = list[id];
            const int l = base + off[id], r = l + splits;
            T.area[l] = res[id].left_area; T.prim_index[l] = T.prim_index[par]; T.bitmap[l] = (int64_t)res[id].left;
            T.area[r] = res[id].right_area; T.prim_index[r] = T.prim_index[par]; T.bitmap[r] = (int64_t)res[id].right;
            T.parent[l] = T.parent[r] = par;
            T.children[2 * par] = l; T.children[2 * par + 1] = r;
        }
        /* PropagateAabbToChildren (lower), LowerTriangleMap.h:602-636 */
        if (prm.propagate_aabbs && splits > 0) {
            for (int id = 0; id < A; ++id) {
                const int par = list[id];
                const int ax = T.info[par] & 3;
                if (ax == LEAF) continue;
                const float s = T.split[par];
                const int l = T.children[2 * par], r = T.children[2 * par + 1];
                const f4 mn = T.bmin[par], mx = T.bmax[par];
                T.bmin[l] = mn;
                T.bmin[r] = {ax == AX_X ? s : mn.x, ax == AX_Y ? s : mn.y, ax == AX_Z ? s : mn.z, 0.0f};
                T.bmax[l] = {ax == AX_X ? s : mx.x, ax == AX_Y ? s : mx.y, ax == AX_Z ? s : mx.z, 0.0f};
                T.bmax[r] = mx;
            }
        }
        list.clear();
        for (int i = 0; i < 2 * splits; ++i) list.push_back(base + i);
        ++level;
    }
    T.lower_levels = level;
    return 0;
}

/* ------------------------------------------------------------------ */
/* Traversal                                                           */
/* ------------------------------------------------------------------ */
struct View {
    const int8_t *info; const float *split; const int *children;
    const int *prim_index; const int64_t *bitmap; const f4 *bmin, *bmax;
    const int *prim_indices; int num_nodes;
    const f4 *w0, *w1, *w2;   /* Woop matrices per triangle (GeometryList::GetWoopValues), or null: MT */
};

/* one triangle: Moller-Trumbore on the vertices or Woop on its matrices */
inline bool intersect(const View &V, const f4 *P0, const f4 *P1, const f4 *P2, int prim, f3 o, f3 d,
                      float &th, float &tu, float &tv) {
    if (V.w0) return woop_test(V.w0[prim], V.w1[prim], V.w2[prim], o, d, th, tu, tv);
    return moller_trumbore(mk3(P0[prim]), mk3(P1[prim]), mk3(P2[prim]), o, d, th, tu, tv);
}

inline orc_hit miss_record() { orc_hit h; h.id = -1; h.t = INFINITY; h.u = 0.0f; h.v = 0.0f; return h; }

/* test the leaf's triangles; returns hit prim or -1 (Raytracer.cu:194-216) */
inline int test_leaf(const View &V, const f4 *P0, const f4 *P1, const f4 *P2, int node, f3 o,
                     f3 d, f3 inv, bool leaf_skip, float &th, float &tu, float &tv, int *cnt) {
    if (cnt) cnt[1]++;
    if (leaf_skip) {
        if (cnt) cnt[3]++;
        if (!ray_box(o, inv, mk3(V.bmin[node]), mk3(V.bmax[node]))) return -1;
    }
    const int pi = V.prim_index[node];
    uint64_t bm = (uint64_t)V.bitmap[node];
    int hit = -1;
    while (bm) {
        const int i = ctz64(bm);
        const int prim = V.prim_indices[pi + i];
        if (cnt) cnt[2]++;
        if (intersect(V, P0, P1, P2, prim, o, d, th, tu, tv)) hit = prim;
        bm &= bm - 1;
    }
    return hit;
}

/* KDRestart / TraceNode, Raytracer.cu:57-106,165-262 (first hit only) */
orc_hit trace_kd_restart(const View &V, const f4 *P0, const f4 *P1, const f4 *P2, f3 o, f3 draw,
                         bool leaf_skip, int root, int *cnt) {
    const f3 d = adjust_dir(draw);
    const f3 inv = {1.0f / d.x, 1.0f / d.y, 1.0f / d.z};
    float th = 0.0f, tu = 0.0f, tv = 0.0f;
    long long visits = 0;
    do {
        if (++visits > (long long)V.num_nodes + 1) break;
        float tNext = F_INF;
        int node = root;
        int8_t axis = V.info[root];
        int steps = 0;
        while ((axis & 3) != LEAF) {
            if (++steps > MAX_DESCENT) return miss_record();
            if (cnt) cnt[0]++;
            const float s = V.split[node];
            const int cl = V.children[2 * node], cr = V.children[2 * node + 1];
            const int a = axis & 3;
            const float ori = comp(o, a), dd = comp(inv, a);
            const float tSplit = (s - ori) * dd;
            if (th < tSplit) {
                node = 0 < dd ? cl : cr;
                tNext = std::fmin(tSplit, tNext);
            } else {
                node = 0 < dd ? cr : cl;
            }
            axis = V.info[node];
        }
        th = tNext * 1.00001f;
        const int hit = test_leaf(V, P0, P1, P2, node, o, d, inv, leaf_skip, th, tu, tv, cnt);
        if (hit != -1) { orc_hit h; h.id = hit; h.t = th; h.u = tu; h.v = tv; return h; }
    } while (th < F_INF);
    return miss_record();
}

/* ShortStackTrace + Stack<N>, ShortStack.cu:58-194, ShortStack.h:43-98 */
orc_hit trace_short_stack(const View &V, const f4 *P0, const f4 *P1, const f4 *P2, f3 o, f3 draw,
                          bool leaf_skip, int S, int root, int *cnt) {
    const f3 d = adjust_dir(draw);
    const f3 inv = {1.0f / d.x, 1.0f / d.y, 1.0f / d.z};
    int stk_node[64]; float stk_t[64];
    int nxt = 0, count = 0;
    float th = 0.0f, tu = 0.0f, tv = 0.0f;
    long long visits = 0;
    do {
        if (++visits > (long long)V.num_nodes + 1) break;
        int node; float tNext;
        if (count == 0) { node = root; tNext = F_INF; }
        else {
            nxt = (nxt == 0 ? S : nxt) - 1;
            count--;
            node = stk_node[nxt]; tNext = stk_t[nxt];
        }
        int8_t info = V.info[node];
        int steps = 0;
        while ((info & 3) != LEAF) {
            if (++steps > MAX_DESCENT) return miss_record();
            if (cnt) cnt[0]++;
            const float s = V.split[node];
            const int cl = V.children[2 * node], cr = V.children[2 * node + 1];
            const int a = info & 3;
            const float ori = comp(o, a), dd = comp(inv, a);
            const float tSplit = (s - ori) * dd;
            const int lower = 0 < dd ? cl : cr;
            const int upper = 0 < dd ? cr : cl;
            if (th < tSplit) {
                node = lower;
                if (tSplit < tNext) {           /* Push, ShortStack.h:57-62 */
                    stk_node[nxt] = upper; stk_t[nxt] = tNext;
                    nxt = nxt == S - 1 ? 0 : nxt + 1;
                    count = count == S ? S : count + 1;
                }
                tNext = std::fmin(tSplit, tNext);
            } else {
                node = upper;
            }
            info = V.info[node];
        }
        th = tNext * 1.00001f;
        const int hit = test_leaf(V, P0, P1, P2, node, o, d, inv, leaf_skip, th, tu, tv, cnt);
        if (hit != -1) { orc_hit h; h.id = hit; h.t = th; h.u = tu; h.v = tv; return h; }
    } while (th < F_INF);
    return miss_record();
}

/* Exhaustive, BruteTracer.cu:54-96 (first hit; no direction adjustment) */
orc_hit trace_brute(const View &V, const f4 *P0, const f4 *P1, const f4 *P2, int ntris, int cap, f3 o, f3 d,
                    int *cnt) {
    const int prims = cap > 0 ? std::min(ntris, cap) : ntris;
    float th = F_INF, tu = 0.0f, tv = 0.0f;
    int hit = -1;
    for (int prim = 0; prim < prims; ++prim) {
        if (intersect(V, P0, P1, P2, prim, o, d, th, tu, tv)) hit = prim;
    }
    if (cnt) cnt[2] += prims;
    if (th < F_INF && hit != -1) { orc_hit h; h.id = hit; h.t = th; h.u = tu; h.v = tv; return h; }
    return miss_record();
}

Params to_params(const orc_build_params *p) {
    Params q;
    q.lower_algorithm = p->lower_algorithm;
    q.empty_space_splitting = p->empty_space_splitting;
    q.empty_space_threshold = p->empty_space_threshold;
    q.split_method = p->split_method;
    q.propagate_aabbs = p->propagate_aabbs;
    q.traversal_cost = p->traversal_cost;
    q.lower_size = p->lower_size;
    q.max_upper_levels = p->max_upper_levels;
    q.sah_leaf_max = p->sah_leaf_max;
    q.sah_empty_bonus = p->sah_empty_bonus;
    q.max_ref_factor = p->max_ref_factor;
    return q;
}

}  // namespace

/* ------------------------------------------------------------------ */
/* Canvas shading: ColorKernels.h:15-85 and the tracers' shading loops */
/* (Raytracer.cu:165-262, ShortStack.cu:96-194, BruteTracer.cu:54-96). */
/* Host-branch arithmetic; the specular power is the correctly rounded */
/* float of the double-precision power.                               */
/* ------------------------------------------------------------------ */
struct Light { f3 pos, ambient, diffuse, specular; };

inline f3 scale(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline f3 add(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline f3 mul(f3 a, f3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline f3 normalize3(f3 v) { return scale(v, 1.0f / std::sqrt(dot(v, v))); }
inline float saturate(float x) { return x > 1.0f ? 1.0f : (x > 0.0f ? x : 0.0f); }
inline float clamp01(float x) { return std::fmax(0.0f, std::fmin(x, 1.0f)); }
inline float spec_pow(float x, float y) { return (float)std::pow((double)x, (double)y); }

struct Color { float r, g, b, a; };

/* PhongLighting, ColorKernels.h:15-45 (shadow = 1: the reference's shadow
 * rays are disabled, Raytracer.cu:219-226) */
inline Color phong(const Color c, f3 normal, f3 point, f3 origin, const Light &L) {
    const f3 lightDir = normalize3(sub(L.pos, point));
    const float diffuse = saturate(dot(lightDir, normal) * 1.0f);
    f3 refl = sub(scale(normal, 2.0f * dot(normal, lightDir)), lightDir);
    refl = normalize3(refl);
    const float stemp = saturate(dot(normalize3(sub(origin, point)), refl));
    const float specProp = 1.0f - c.a;
    const float specular = specProp * spec_pow(stemp, 128.0f * specProp);
    const f3 base = {c.r, c.g, c.b};
    const f3 light = add(mul(base, add(L.ambient, scale(L.diffuse, diffuse))),
                         scale(scale(scale(L.specular, specular), 10.0f), 1.0f));
    const float alpha = c.a < specular ? specular : c.a;
    return {clamp01(light.x), clamp01(light.y), clamp01(light.z), alpha};
}

/* Lighting, ColorKernels.h:47-77: shade the hit and turn the ray into its
 * mirror reflection from the hit point */
inline Color lighting(const orc_hit &h, f3 &origin, f3 &dir, const f4 *n0, const f4 *n1, const f4 *n2,
                      const uint8_t *c0, const Light &L) {
    const f3 point = add(origin, scale(dir, h.t));
    const float w = 1.0f - h.u - h.v;
    f3 normal = add(add(scale(mk3(n0[h.id]), w), scale(mk3(n1[h.id]), h.u)), scale(mk3(n2[h.id]), h.v));
    normal = normalize3(normal);
    const uint8_t *cc = c0 + 4 * (size_t)h.id;
    const Color col = {cc[0] / 255.0f, cc[1] / 255.0f, cc[2] / 255.0f, cc[3] / 255.0f};
    dir = add(scale(normal, 2.0f * dot(normal, scale(dir, -1.0f))), dir);
    dir = normalize3(dir);
    const Color out = phong(col, normal, point, origin, L);
    origin = add(point, scale(dir, 0.00001f));
    return out;
}

/* BlendColor, ColorKernels.h:79-84 */
inline Color blend(const Color o, Color n) {
    const float a = n.a;
    n.a = 1.0f;
    const float k = (1.0f - o.a) * a;
    return {o.r + k * n.r, o.g + k * n.g, o.b + k * n.b, o.a + k * n.a};
}

struct orc_tree { Tree T; };

extern "C" {

const char *orc_last_error(void) { return g_err.c_str(); }

int orc_build_batched(const float *p0, const float *p1, const float *p2, int n,
                      const int *mesh_offsets, int num_meshes, const orc_build_params *params,
                      orc_tree **out) {
    g_err.clear();
    if (!out || !params || n < 0 || (n > 0 && (!p0 || !p1 || !p2))) { g_err = "invalid argument"; return 1; }
    if (num_meshes < 1 || !mesh_offsets) { g_err = "need at least one mesh and its offsets"; return 1; }
    for (int m = 0; m < num_meshes; ++m)
        if (mesh_offsets[m] > mesh_offsets[m + 1]) { g_err = "mesh offsets must be non-decreasing"; return 1; }
    if (mesh_offsets[0] != 0 || mesh_offsets[num_meshes] != n) { g_err = "mesh offsets must span [0, n]"; return 1; }
    const Params prm = to_params(params);
    if (prm.split_method == 2) { g_err = "split method not implemented"; return 2; }
    if (prm.split_method != 0 && prm.split_method != 1) { g_err = "unknown split method"; return 1; }
    if (prm.lower_size < 1 || prm.lower_size > 64) { g_err = "lower_size must be in [1, 64]"; return 1; }
    if (prm.lower_algorithm < 0 || prm.lower_algorithm > 2) { g_err = "unknown lower algorithm"; return 1; }
    orc_tree *t = new orc_tree();
    const f4 *P0 = reinterpret_cast<const f4 *>(p0), *P1 = reinterpret_cast<const f4 *>(p1),
             *P2 = reinterpret_cast<const f4 *>(p2);
    std::vector<f4> pmin, pmax;
    int rc = build_upper(P0, P1, P2, n, mesh_offsets, num_meshes, prm, t->T, pmin, pmax);
    if (!rc) rc = build_lower(P0, P1, P2, prm, t->T, pmin, pmax);
    if (rc) { delete t; return rc; }
    *out = t;
    return 0;
}

int orc_build(const float *p0, const float *p1, const float *p2, int n,
              const orc_build_params *params, orc_tree **out) {
    const int off[2] = {0, n < 0 ? 0 : n};
    return orc_build_batched(p0, p1, p2, n, off, 1, params, out);
}

int orc_num_roots(const orc_tree *t) { return t->T.num_roots; }

void orc_woop_matrix(const float *a, const float *b, const float *c, float *m) {
    f4 m0, m1, m2;
    woop_matrix({a[0], a[1], a[2]}, {b[0], b[1], b[2]}, {c[0], c[1], c[2]}, m0, m1, m2);
    const f4 ms[3] = {m0, m1, m2};
    for (int i = 0; i < 3; ++i) {
        m[4 * i] = ms[i].x; m[4 * i + 1] = ms[i].y; m[4 * i + 2] = ms[i].z; m[4 * i + 3] = ms[i].w;
    }
}

int orc_balanced_node(int n, const float *bmin, const float *bmax, long long bitmap, float traversal_cost,
                      int lower_size, int *axis, float *pos, long long *left, long long *right, int *cost) {
    if (n < 0 || n > 64) { g_err = "n must be in [0, 64]"; return -1; }
    f3 mins[64], maxs[64];
    for (int i = 0; i < n; ++i) {
        mins[i] = {bmin[3 * i], bmin[3 * i + 1], bmin[3 * i + 2]};
        maxs[i] = {bmax[3 * i], bmax[3 * i + 1], bmax[3 * i + 2]};
    }
    const BalancedResult r = balanced_node(mins, maxs, (uint64_t)bitmap, traversal_cost, lower_size);
    *axis = r.axis; *pos = r.pos; *left = (long long)r.left; *right = (long long)r.right; *cost = r.cost;
    return r.split;
}

int orc_tri_box_step3(const float *v0, const float *v1, const float *v2, const float *bmin, const float *bmax) {
    return tri_box_step3({v0[0], v0[1], v0[2]}, {v1[0], v1[1], v1[2]}, {v2[0], v2[1], v2[2]},
                         {bmin[0], bmin[1], bmin[2]}, {bmax[0], bmax[1], bmax[2]});
}

void orc_free(orc_tree *t) { delete t; }
int orc_num_nodes(const orc_tree *t) { return t->T.size(); }
int orc_num_prim_indices(const orc_tree *t) { return (int)t->T.prim_indices.size(); }
int orc_num_leaf_ids(const orc_tree *t) { return (int)t->T.leaf_ids.size(); }
int orc_upper_levels(const orc_tree *t) { return t->T.upper_levels; }
int orc_lower_levels(const orc_tree *t) { return t->T.lower_levels; }
long long orc_upper_level_refs(const orc_tree *t, int l) {
    return (l >= 0 && l < (int)t->T.level_refs.size()) ? t->T.level_refs[l] : -1;
}
long long orc_lower_level_nodes(const orc_tree *t, int l) {
    return (l >= 0 && l < (int)t->T.lower_level_nodes.size()) ? t->T.lower_level_nodes[l] : -1;
}

void orc_copy_tree(const orc_tree *t, int8_t *info, float *split_pos, float *aabb_min,
                   float *aabb_max, int *primitive_index, long long *primitive_bitmap,
                   int *children, float *surface_area, int *parent, int *prim_indices,
                   int *leaf_ids) {
    const Tree &T = t->T;
    const size_t n = (size_t)T.size();
    if (info) std::memcpy(info, T.info.data(), n);
    if (split_pos) std::memcpy(split_pos, T.split.data(), n * 4);
    if (aabb_min) std::memcpy(aabb_min, T.bmin.data(), n * 16);
    if (aabb_max) std::memcpy(aabb_max, T.bmax.data(), n * 16);
    if (primitive_index) std::memcpy(primitive_index, T.prim_index.data(), n * 4);
    if (primitive_bitmap) std::memcpy(primitive_bitmap, T.bitmap.data(), n * 8);
    if (children) std::memcpy(children, T.children.data(), n * 8);
    if (surface_area) std::memcpy(surface_area, T.area.data(), n * 4);
    if (parent) std::memcpy(parent, T.parent.data(), n * 4);
    if (prim_indices) std::memcpy(prim_indices, T.prim_indices.data(), T.prim_indices.size() * 4);
    if (leaf_ids) std::memcpy(leaf_ids, T.leaf_ids.data(), T.leaf_ids.size() * 4);
}

int orc_tree_from_arrays(int num_nodes, const int8_t *info, const float *split_pos,
                         const float *aabb_min, const float *aabb_max, const int *primitive_index,
                         const long long *primitive_bitmap, const int *children,
                         int num_prim_indices, const int *prim_indices, orc_tree **out) {
    if (num_nodes < 1 || !out) { g_err = "invalid argument"; return 1; }
    orc_tree *t = new orc_tree();
    Tree &T = t->T;
    T.resize(num_nodes);
    std::memcpy(T.info.data(), info, num_nodes);
    std::memcpy(T.split.data(), split_pos, (size_t)num_nodes * 4);
    std::memcpy(T.bmin.data(), aabb_min, (size_t)num_nodes * 16);
    std::memcpy(T.bmax.data(), aabb_max, (size_t)num_nodes * 16);
    std::memcpy(T.prim_index.data(), primitive_index, (size_t)num_nodes * 4);
    std::memcpy(T.bitmap.data(), primitive_bitmap, (size_t)num_nodes * 8);
    std::memcpy(T.children.data(), children, (size_t)num_nodes * 8);
    T.prim_indices.assign(prim_indices, prim_indices + num_prim_indices);
    *out = t;
    return 0;
}

int orc_stats_of(const orc_tree *t, orc_stats *s) {
    const Tree &T = t->T;
    std::memset(s, 0, sizeof(*s));
    s->num_nodes = T.size();
    s->upper_levels = T.upper_levels;
    s->lower_levels = T.lower_levels;
    s->num_prim_indices = (int)T.prim_indices.size();
    s->num_upper_leaves = (int)T.leaf_ids.size();
    if (T.size() == 0) return 0;
    std::vector<std::pair<int, int>> stack;
    for (int r = T.num_roots - 1; r >= 0; --r) stack.push_back({r, 0});
    double depth_sum = 0.0;
    long long guard = 0;
    while (!stack.empty()) {
        if (++guard > 4LL * T.size() + 4) { g_err = "tree is not acyclic"; return 1; }
        auto [node, depth] = stack.back();
        stack.pop_back();
        if (node < 0 || node >= T.size()) { g_err = "child index out of range"; return 1; }
        s->reachable_nodes++;
        s->max_depth = std::max(s->max_depth, depth);
        if ((T.info[node] & 3) == LEAF) {
            s->leaves++;
            const int c = popc64((uint64_t)T.bitmap[node]);
            if (c == 0) s->empty_leaves++;
            else depth_sum += depth;
            s->leaf_refs += c;
            s->max_leaf_size = std::max(s->max_leaf_size, c);
        } else {
            s->inner_nodes++;
            stack.push_back({T.children[2 * node + 1], depth + 1});
            stack.push_back({T.children[2 * node], depth + 1});
        }
    }
    const int nonempty = s->leaves - s->empty_leaves;
    s->mean_leaf_depth = nonempty ? depth_sum / nonempty : 0.0;
    return 0;
}

int orc_trace_batched(const orc_tree *t, const float *p0, const float *p1, const float *p2, int ntris,
                      int algo, const float *org, const float *dir, const int *ray_root, int nrays,
                      const orc_trace_params *params, orc_hit *out, int *counters) {
    g_err.clear();
    if (nrays < 0 || !out || !params) { g_err = "invalid argument"; return 1; }
    if (algo != 0 && !t) { g_err = "tree required"; return 1; }
    if (ray_root && algo == 0) { g_err = "brute force has no roots"; return 1; }
    const int roots = t ? t->T.num_roots : 1;
    const f4 *P0 = reinterpret_cast<const f4 *>(p0), *P1 = reinterpret_cast<const f4 *>(p1),
             *P2 = reinterpret_cast<const f4 *>(p2);
    const f4 *O = reinterpret_cast<const f4 *>(org), *D = reinterpret_cast<const f4 *>(dir);
    View V{};
    if (t) {
        const Tree &T = t->T;
        V = View{T.info.data(), T.split.data(), T.children.data(), T.prim_index.data(),
                 T.bitmap.data(), T.bmin.data(), T.bmax.data(), T.prim_indices.data(), T.size(),
                 nullptr, nullptr, nullptr};
    }
    /* CreateWoopValues (GeometryList.cu:80-101): matrices of (p1, p2, p0) */
    std::vector<f4> w0, w1, w2;
    if (params->intersection == 1) {
        w0.resize((size_t)std::max(ntris, 0)); w1.resize(w0.size()); w2.resize(w0.size());
#pragma omp parallel for schedule(static)
        for (int k = 0; k < ntris; ++k) woop_matrix(mk3(P1[k]), mk3(P2[k]), mk3(P0[k]), w0[k], w1[k], w2[k]);
        V.w0 = w0.data(); V.w1 = w1.data(); V.w2 = w2.data();
    } else if (params->intersection != 0) {
        g_err = "unknown intersection algorithm";
        return 1;
    }
    const int S = params->stack_size > 0 ? params->stack_size : 4;
    if (S > 64) { g_err = "stack_size must be <= 64"; return 1; }
    const bool skip = params->leaf_skipping != 0;
#ifdef _OPENMP
    const int nt = params->threads > 0 ? params->threads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 256) num_threads(nt)
#endif
    for (int r = 0; r < nrays; ++r) {
        int *cnt = counters ? counters + 4 * (size_t)r : nullptr;
        if (cnt) cnt[0] = cnt[1] = cnt[2] = cnt[3] = 0;
        const f3 o = mk3(O[r]), d = mk3(D[r]);
        const int root = ray_root ? ray_root[r] : 0;
        if (root < 0 || root >= roots) { out[r] = miss_record(); continue; }   /* no such mesh */
        if (algo == 0) out[r] = trace_brute(V, P0, P1, P2, ntris, params->brute_cap, o, d, cnt);
        else if (algo == 1) out[r] = trace_kd_restart(V, P0, P1, P2, o, d, skip, root, cnt);
        else out[r] = trace_short_stack(V, P0, P1, P2, o, d, skip, S, root, cnt);
    }
    return 0;
}

int orc_trace_canvas(const orc_tree *t, const float *p0, const float *p1, const float *p2, int ntris,
                     int algo, const float *org, const float *dir, int nrays, const float *n0,
                     const float *n1, const float *n2, const unsigned char *c0, const float *light,
                     int max_bounces, const orc_trace_params *params, unsigned char *canvas) {
    g_err.clear();
    if (nrays < 0 || !canvas || !params || !light || (nrays > 0 && (!n0 || !n1 || !n2 || !c0))) {
        g_err = "invalid argument";
        return 1;
    }
    if (algo != 0 && !t) { g_err = "tree required"; return 1; }
    const f4 *P0 = reinterpret_cast<const f4 *>(p0), *P1 = reinterpret_cast<const f4 *>(p1),
             *P2 = reinterpret_cast<const f4 *>(p2);
    const f4 *O = reinterpret_cast<const f4 *>(org), *D = reinterpret_cast<const f4 *>(dir);
    const f4 *N0 = reinterpret_cast<const f4 *>(n0), *N1 = reinterpret_cast<const f4 *>(n1),
             *N2 = reinterpret_cast<const f4 *>(n2);
    const Light L = {{light[0], light[1], light[2]}, {light[3], light[4], light[5]},
                     {light[6], light[7], light[8]}, {light[9], light[10], light[11]}};
    View V{};
    if (t) {
        const Tree &T = t->T;
        V = View{T.info.data(), T.split.data(), T.children.data(), T.prim_index.data(),
                 T.bitmap.data(), T.bmin.data(), T.bmax.data(), T.prim_indices.data(), T.size(),
                 nullptr, nullptr, nullptr};
    }
    std::vector<f4> w0, w1, w2;
    if (params->intersection == 1) {
        w0.resize((size_t)std::max(ntris, 0)); w1.resize(w0.size()); w2.resize(w0.size());
        for (int k = 0; k < ntris; ++k) woop_matrix(mk3(P1[k]), mk3(P2[k]), mk3(P0[k]), w0[k], w1[k], w2[k]);
        V.w0 = w0.data(); V.w1 = w1.data(); V.w2 = w2.data();
    }
    const int S = params->stack_size > 0 ? params->stack_size : 4;
    const bool skip = params->leaf_skipping != 0;
#ifdef _OPENMP
    const int nt = params->threads > 0 ? params->threads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 256) num_threads(nt)
#endif
    for (int r = 0; r < nrays; ++r) {
        f3 o = mk3(O[r]), d = mk3(D[r]);
        if (algo != 0) d = adjust_dir(d);
        Color col = {0.0f, 0.0f, 0.0f, 0.0f};
        for (int b = 0; b < max_bounces; ++b) {
            orc_hit h;
            if (algo == 0) h = trace_brute(V, P0, P1, P2, ntris, params->brute_cap, o, d, nullptr);
            else if (algo == 1) h = trace_kd_restart(V, P0, P1, P2, o, d, skip, 0, nullptr);
            else h = trace_short_stack(V, P0, P1, P2, o, d, skip, S, 0, nullptr);
            if (h.id < 0) break;
            col = blend(col, lighting(h, o, d, N0, N1, N2, c0, L));
            if (algo != 0) d = adjust_dir(d);
            if (!(col.a < 0.97f)) break;
        }
        unsigned char *px = canvas + 4 * (size_t)r;
        px[0] = (unsigned char)(col.r * 255.0f); px[1] = (unsigned char)(col.g * 255.0f);
        px[2] = (unsigned char)(col.b * 255.0f); px[3] = (unsigned char)(col.a * 255.0f);
    }
    return 0;
}

int orc_trace(const orc_tree *t, const float *p0, const float *p1, const float *p2, int ntris,
              int algo, const float *org, const float *dir, int nrays,
              const orc_trace_params *params, orc_hit *out, int *counters) {
    return orc_trace_batched(t, p0, p1, p2, ntris, algo, org, dir, nullptr, nrays, params, out, counters);
}

int orc_mt32(const float *v0, const float *v1, const float *v2, const float *o, const float *d,
             float t_max, float *tuv) {
    float th = t_max, tu = 0.0f, tv = 0.0f;
    const bool hit = moller_trumbore({v0[0], v0[1], v0[2]}, {v1[0], v1[1], v1[2]},
                                     {v2[0], v2[1], v2[2]}, {o[0], o[1], o[2]},
                                     {d[0], d[1], d[2]}, th, tu, tv);
    tuv[0] = th; tuv[1] = tu; tuv[2] = tv;
    return hit ? 1 : 0;
}

void orc_mt64(const double *v0, const double *v1, const double *v2, const double *o,
              const double *d, double *r) {
    const double e1[3] = {v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2]};
    const double e2[3] = {v2[0] - v0[0], v2[1] - v0[1], v2[2] - v0[2]};
    const double p[3] = {d[1] * e2[2] - d[2] * e2[1], d[2] * e2[0] - d[0] * e2[2], d[0] * e2[1] - d[1] * e2[0]};
    const double T[3] = {o[0] - v0[0], o[1] - v0[1], o[2] - v0[2]};
    const double q[3] = {T[1] * e1[2] - T[2] * e1[1], T[2] * e1[0] - T[0] * e1[2], T[0] * e1[1] - T[1] * e1[0]};
    const double det = p[0] * e1[0] + p[1] * e1[1] + p[2] * e1[2];
    const double inv = 1.0 / det;
    r[0] = inv * (q[0] * e2[0] + q[1] * e2[1] + q[2] * e2[2]);
    r[1] = inv * (p[0] * T[0] + p[1] * T[1] + p[2] * T[2]);
    r[2] = inv * (q[0] * d[0] + q[1] * d[1] + q[2] * d[2]);
    r[3] = det;
}

static inline void cross3(const double *a, const double *b, double *c) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

void orc_mt64_adjoint(const double *v0, const double *v1, const double *v2, const double *o,
                      const double *d, double gt, double gu, double gv, double *g_v0,
                      double *g_v1, double *g_v2, double *g_o, double *g_d) {
    double e1[3], e2[3], T[3], p[3], q[3];
    for (int i = 0; i < 3; ++i) { e1[i] = v1[i] - v0[i]; e2[i] = v2[i] - v0[i]; T[i] = o[i] - v0[i]; }
    cross3(d, e2, p);
    cross3(T, e1, q);
    const double det = p[0] * e1[0] + p[1] * e1[1] + p[2] * e1[2];
    const double inv = 1.0 / det;
    const double a = q[0] * e2[0] + q[1] * e2[1] + q[2] * e2[2];
    const double b = p[0] * T[0] + p[1] * T[1] + p[2] * T[2];
    const double c = q[0] * d[0] + q[1] * d[1] + q[2] * d[2];
    const double ga = gt * inv, gb = gu * inv, gc = gv * inv;
    const double ginv = gt * a + gu * b + gv * c;
    const double gdet = -ginv * inv * inv;
    double gq[3], gp[3], gT[3], ge1[3], ge2[3], gd[3];
    for (int i = 0; i < 3; ++i) {
        gq[i] = ga * e2[i] + gc * d[i];
        ge2[i] = ga * q[i];
        gp[i] = gb * T[i] + gdet * e1[i];
        gT[i] = gb * p[i];
        gd[i] = gc * q[i];
        ge1[i] = gdet * p[i];
    }
    double tmp[3];
    cross3(e1, gq, tmp);            /* q = T x e1: dT += e1 x gq */
    for (int i = 0; i < 3; ++i) gT[i] += tmp[i];
    cross3(gq, T, tmp);             /* de1 += gq x T */
    for (int i = 0; i < 3; ++i) ge1[i] += tmp[i];
    cross3(e2, gp, tmp);            /* p = d x e2: dd += e2 x gp */
    for (int i = 0; i < 3; ++i) gd[i] += tmp[i];
    cross3(gp, d, tmp);             /* de2 += gp x d */
    for (int i = 0; i < 3; ++i) ge2[i] += tmp[i];
    for (int i = 0; i < 3; ++i) {
        g_o[i] = gT[i];
        g_d[i] = gd[i];
        g_v1[i] = ge1[i];
        g_v2[i] = ge2[i];
        g_v0[i] = -gT[i] - ge1[i] - ge2[i];
    }
}

int orc_sah_node(int n, const float *bmin, const float *bmax, const float *areas, long long bitmap,
                 float node_area, float traversal_cost, int sah_leaf_max, float sah_empty_bonus,
                 int *axis, float *pos, long long *left, long long *right, float *left_area,
                 float *right_area, float *cost) {
    if (n < 0 || n > 64) return -1;
    f3 mins[64], maxs[64];
    for (int i = 0; i < n; ++i) {
        mins[i] = {bmin[3 * i], bmin[3 * i + 1], bmin[3 * i + 2]};
        maxs[i] = {bmax[3 * i], bmax[3 * i + 1], bmax[3 * i + 2]};
    }
    const uint64_t bmp = (uint64_t)bitmap & (n == 64 ? ~(uint64_t)0 : (((uint64_t)1 << n) - 1));
    SahResult r = sah_node(mins, maxs, areas, bmp, node_area, traversal_cost, sah_leaf_max, sah_empty_bonus);
    *axis = r.axis; *pos = r.pos; *left = (long long)r.left; *right = (long long)r.right;
    *left_area = r.left_area; *right_area = r.right_area; *cost = r.cost;
    return r.split ? 1 : 0;
}

float orc_tri_area(const float *v0, const float *v1, const float *v2) {
    return tri_area({v0[0], v0[1], v0[2]}, {v1[0], v1[1], v1[2]}, {v2[0], v2[1], v2[2]});
}

int orc_trace_backward32(const float *p0, const float *p1, const float *p2, int ntris,
                         const float *org, const float *dir, const orc_hit *hits, int nrays,
                         const float *grad_t, const float *grad_u, const float *grad_v,
                         const int *faces, int num_vertices, float *grad_vertices,
                         float *grad_org, float *grad_dir) {
    g_err.clear();
    if (nrays < 0 || num_vertices < 0) { g_err = "invalid argument"; return 1; }
    const f4 *P0 = reinterpret_cast<const f4 *>(p0), *P1 = reinterpret_cast<const f4 *>(p1),
             *P2 = reinterpret_cast<const f4 *>(p2);
    const f4 *O = reinterpret_cast<const f4 *>(org), *D = reinterpret_cast<const f4 *>(dir);
    std::vector<float> corner(9 * (size_t)nrays, 0.0f);
    std::vector<int> key(3 * (size_t)nrays, -1);
    for (int r = 0; r < nrays; ++r) {
        const int id = hits[r].id;
        f3 g_o = {0, 0, 0}, g_d = {0, 0, 0}, g[3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        if (id >= 0 && id < ntris) {
            const float gt = grad_t ? grad_t[r] : 0.0f, gu = grad_u ? grad_u[r] : 0.0f,
                        gv = grad_v ? grad_v[r] : 0.0f;
            const f3 v0 = mk3(P0[id]), v1 = mk3(P1[id]), v2 = mk3(P2[id]);
            const f3 o = mk3(O[r]), d = mk3(D[r]);
            const f3 e1 = sub(v1, v0), e2 = sub(v2, v0), T = sub(o, v0);
            const f3 p = cross(d, e2), q = cross(T, e1);
            const float det = dot(p, e1);
            const float inv = 1.0f / det;
            const float a = dot(q, e2), b = dot(p, T), c = dot(q, d);
            const float ga = gt * inv, gb = gu * inv, gc = gv * inv;
            const float ginv = gt * a + gu * b + gv * c;
            const float gdet = -ginv * inv * inv;
            auto sc = [](f3 x, float k) { return f3{x.x * k, x.y * k, x.z * k}; };
            auto ad = [](f3 x, f3 y) { return f3{x.x + y.x, x.y + y.y, x.z + y.z}; };
            const f3 gq = ad(sc(e2, ga), sc(d, gc));
            const f3 gp = ad(sc(T, gb), sc(e1, gdet));
            const f3 gT = ad(sc(p, gb), cross(e1, gq));
            const f3 ge1 = ad(sc(p, gdet), cross(gq, T));
            const f3 gd = ad(sc(q, gc), cross(e2, gp));
            const f3 ge2 = ad(sc(q, ga), cross(gp, d));
            g_o = gT; g_d = gd; g[1] = ge1; g[2] = ge2;
            g[0] = {-gT.x - ge1.x - ge2.x, -gT.y - ge1.y - ge2.y, -gT.z - ge1.z - ge2.z};
            for (int k = 0; k < 3; ++k) key[3 * (size_t)r + k] = faces ? faces[3 * id + k] : 3 * id + k;
        }
        for (int k = 0; k < 3; ++k) {
            corner[9 * (size_t)r + 3 * k] = g[k].x;
            corner[9 * (size_t)r + 3 * k + 1] = g[k].y;
            corner[9 * (size_t)r + 3 * k + 2] = g[k].z;
        }
        if (grad_org) { grad_org[3 * r] = g_o.x; grad_org[3 * r + 1] = g_o.y; grad_org[3 * r + 2] = g_o.z; }
        if (grad_dir) { grad_dir[3 * r] = g_d.x; grad_dir[3 * r + 1] = g_d.y; grad_dir[3 * r + 2] = g_d.z; }
    }
    std::fill(grad_vertices, grad_vertices + 3 * (size_t)num_vertices, 0.0f);
    for (size_t e = 0; e < key.size(); ++e) {       /* (ray, corner) order */
        const int v = key[e];
        if (v < 0) continue;
        if (v >= num_vertices) { g_err = "vertex id out of range"; return 1; }
        for (int c = 0; c < 3; ++c) grad_vertices[3 * (size_t)v + c] += corner[3 * e + c];
    }
    return 0;
}

int orc_trace_backward64(const float *p0, const float *p1, const float *p2, int ntris,
                         const float *org, const float *dir, const orc_hit *hits, int nrays,
                         const float *grad_t, const float *grad_u, const float *grad_v,
                         const int *faces, int num_vertices, double *grad_vertices,
                         double *grad_org, double *grad_dir) {
    g_err.clear();
    if (nrays < 0 || num_vertices < 0) { g_err = "invalid argument"; return 1; }
    std::fill(grad_vertices, grad_vertices + 3 * (size_t)num_vertices, 0.0);
    for (int r = 0; r < nrays; ++r) {
        const int id = hits[r].id;
        double go[3] = {0, 0, 0}, gd[3] = {0, 0, 0};
        if (id >= 0 && id < ntris) {
            double v[3][3], o[3], d[3], gv3[3][3];
            for (int c = 0; c < 3; ++c) {
                v[0][c] = p0[4 * (size_t)id + c]; v[1][c] = p1[4 * (size_t)id + c]; v[2][c] = p2[4 * (size_t)id + c];
                o[c] = org[4 * (size_t)r + c]; d[c] = dir[4 * (size_t)r + c];
            }
            orc_mt64_adjoint(v[0], v[1], v[2], o, d, grad_t ? grad_t[r] : 0.0, grad_u ? grad_u[r] : 0.0,
                             grad_v ? grad_v[r] : 0.0, gv3[0], gv3[1], gv3[2], go, gd);
            for (int k = 0; k < 3; ++k) {
                const int vid = faces ? faces[3 * id + k] : 3 * id + k;
                if (vid < 0 || vid >= num_vertices) { g_err = "vertex id out of range"; return 1; }
                for (int c = 0; c < 3; ++c) grad_vertices[3 * (size_t)vid + c] += gv3[k][c];
            }
        }
        if (grad_org) for (int c = 0; c < 3; ++c) grad_org[3 * r + c] = go[c];
        if (grad_dir) for (int c = 0; c < 3; ++c) grad_dir[3 * r + c] = gd[c];
    }
    return 0;
}

/* CreateRays + Unproject, IRayTracer.cu:45-70 (host branch: IEEE division,
 * normalize = v * (1.0f / sqrtf(dot(v, v)))), with the optional seeded
 * sub-pixel jitter: PCG32 (O'Neill 2014, pcg32_srandom_r/pcg32_random_r)
 * stream (seed, 1 << 40 | id), two 24-bit uniforms. */
void orc_create_rays(const float *m, const float *cam, int width, int height, unsigned long long seed,
                     int jitter, float *org, float *dir) {
    const long long n = (long long)width * height;
    const unsigned long long MULT = 6364136223846793005ull;
    for (long long id = 0; id < n; ++id) {
        const int x = (int)(id % width), y = (int)(id / width);
        float jx = 0.0f, jy = 0.0f;
        if (jitter) {
            const unsigned long long inc = (((1ull << 40) | (unsigned long long)id) << 1) | 1ull;
            unsigned long long state = 0ull * MULT + inc;
            state += seed;
            state = state * MULT + inc;
            float u[2];
            for (int k = 0; k < 2; ++k) {
                const unsigned long long old = state;
                state = old * MULT + inc;
                const uint32_t xs = (uint32_t)(((old >> 18) ^ old) >> 27);
                const uint32_t rot = (uint32_t)(old >> 59);
                const uint32_t r = (xs >> rot) | (xs << ((32u - rot) & 31u));
                u[k] = (float)(r >> 8) * (1.0f / 16777216.0f);
            }
            jx = u[0]; jy = u[1];
        }
        const float sx = (((float)x + jx) / (float)width) * 2.0f - 1.0f;
        const float sy = (((float)y + jy) / (float)height) * 2.0f - 1.0f;
        const float c0 = m[0] * sx + m[4] * sy + m[8] + m[12];
        const float c1 = m[1] * sx + m[5] * sy + m[9] + m[13];
        const float c2 = m[2] * sx + m[6] * sy + m[10] + m[14];
        const float c3 = m[3] * sx + m[7] * sy + m[11] + m[15];
        const f3 v = sub({c0 / c3, c1 / c3, c2 / c3}, {cam[0], cam[1], cam[2]});
        const float inv = 1.0f / std::sqrt(dot(v, v));
        float *o = org + 4 * id, *d = dir + 4 * id;
        o[0] = cam[0]; o[1] = cam[1]; o[2] = cam[2]; o[3] = 1.0f;
        d[0] = v.x * inv; d[1] = v.y * inv; d[2] = v.z * inv; d[3] = 0.0f;
    }
}

}  // extern "C"
