// pm_trace.hip -- kd-restart, short-stack and brute-force traversal for
// gfx950 (wave64).  One lane per ray; the short stack lives in LDS,
// interleaved by lane so a wave's 64 accesses hit 64 different banks.
//
// Semantics (first hit of each ray):
//   kd-restart   Raytracer.cu:57-106 (TraceNode), :165-262 (KDRestart)
//   short-stack  ShortStack.cu:58-94 (TraceNode), :96-194, ShortStack.h:43-98
//   brute force  BruteTracer.cu:54-96 (Exhaustive)
//
// The tracers read a trace view built once per tree (build_trace_view):
//   packed[node]   int4 {split bits | prim index, left | bitmap lo,
//                        right | bitmap hi, axis}: one 16-byte load per visit
//   leaf_tris[3k]  float4 v0 (w = triangle id bits), e1 = v1 - v0,
//                  e2 = v2 - v0 for leaf slot k: a leaf's triangles are
//                  contiguous and need no prim-index indirection.
// e1/e2 are the same IEEE subtractions Moller-Trumbore performs, so hits are
// bit-identical to testing the original vertices.
#include <algorithm>
#include <climits>

#include "pm_device.h"
#include "pm_internal.h"


namespace pmkd {

namespace {

constexpr int TB = 256;   // threads per block (4 waves)

__device__ __forceinline__ pm_hit miss_record() {
    pm_hit h;
    h.id = -1; h.t = __int_as_float(0x7f800000); h.u = 0.0f; h.v = 0.0f;
    return h;
}

// Moller-Trumbore (IRayTracer.h:111-145) from precomputed edges.  The same
// values (t = inv (q.e2), u = inv (p.T), v = inv (q.d), inv = 1/det) and the
// same acceptance (0 <= t <= th, u >= 0, v >= 0, u + v <= 1), tested in the
// order that rejects soonest: u first (u <= 1 is implied by v >= 0 and
// u + v <= 1, since rounding is monotonic), then v, then t.  Rays of a wave
// that test the same triangle mostly agree, so the wave skips the rest.
__device__ __forceinline__ bool mt_edges(f3 v0, f3 e1, f3 e2, f3 o, f3 d, float &th, float &tu,
                                         float &tv) {
    const f3 p = cross3(d, e2);
    const f3 t = sub3(o, v0);
    const float inv = rcp_rn(dot3(p, e1));
    const float hy = inv * dot3(p, t);
    if (!(hy >= 0.0f && hy <= 1.0f)) return false;
    const f3 q = cross3(t, e1);
    const float hz = inv * dot3(q, d);
    if (!(hz >= 0.0f && hy + hz <= 1.0f)) return false;
    const float hx = inv * dot3(q, e2);
    if (!(0.0f <= hx && hx <= th)) return false;
    th = hx; tu = hy; tv = hz;
    return true;
}

// Leaf test (Raytracer.cu:194-216): optional box test, then Moller-Trumbore
// over the leaf bitmap in ascending bit order; the next triangle's data is
// loaded before the current one is tested.
template <bool SKIP, bool COUNT, bool WOOP>
__device__ __forceinline__ int test_leaf(const TreeDev &T, const int4 nd, int node, f3 o, f3 d,
                                         f3 inv, float &th, float &tu, float &tv, int4 &cnt) {
    if (COUNT) cnt.y++;
    if (SKIP) {
        if (COUNT) cnt.w++;
        if (!ray_box(o, inv, mk3(T.bmin[node]), mk3(T.bmax[node]))) return -1;
    }
    unsigned long long bm = ((unsigned long long)(unsigned)nd.z << 32) | (unsigned)nd.y;
    if (!bm) return -1;
    if (WOOP) {   // the slot's Woop matrices; the id comes from the prim indices
        const float4 *W = T.leaf_woop + 3ll * nd.x;
        int slot = -1;
        for (; bm; bm &= bm - 1) {
            const int i = __builtin_ctzll(bm);
            if (COUNT) cnt.z++;
            if (woop_test(W[3 * i], W[3 * i + 1], W[3 * i + 2], o, d, th, tu, tv)) slot = i;
        }
        return slot < 0 ? -1 : T.prim_indices[nd.x + slot];
    }
    // two register sets, A and B, alternate: the next triangle loads while
    // the current one is tested.  The loads are unconditional (past the last
    // triangle they re-read the first one) so that a test waits for its own
    // loads only.
    const float4 *L = T.leaf_tris + 3ll * nd.x;
    const int first = __builtin_ctzll(bm);
    bm &= bm - 1;
    float4 a0 = L[3 * first], a1 = L[3 * first + 1], a2 = L[3 * first + 2];
    int hit = -1;
    for (;;) {
        const bool more_b = bm != 0;
        const int j = more_b ? __builtin_ctzll(bm) : first;
        bm &= bm - 1;
        const float4 b0 = L[3 * j], b1 = L[3 * j + 1], b2 = L[3 * j + 2];
        __builtin_amdgcn_sched_barrier(0);         // issue the loads before the test
        if (COUNT) cnt.z++;
        if (mt_edges(mk3(a0), mk3(a1), mk3(a2), o, d, th, tu, tv)) hit = __float_as_int(a0.w);
        if (!more_b) break;
        const bool more_a = bm != 0;
        const int k = more_a ? __builtin_ctzll(bm) : first;
        bm &= bm - 1;
        a0 = L[3 * k]; a1 = L[3 * k + 1]; a2 = L[3 * k + 2];
        __builtin_amdgcn_sched_barrier(0);
        if (COUNT) cnt.z++;
        if (mt_edges(mk3(b0), mk3(b1), mk3(b2), o, d, th, tu, tv)) hit = __float_as_int(b0.w);
        if (!more_a) break;
    }
    return hit;
}

// Traversal state of one ray between passes of the outer loop.
struct RayState {
    f3 o, d, inv;
    float th, tu, tv;
    int root_id;
    int nxt, count;            // short stack: ring position (slot * TB) and fill
    int visits;
};

__device__ __forceinline__ RayState ray_start(const TreeDev &T, int root_id, f3 o, f3 d) {
    RayState R;
    R.o = o;
    R.d = d;
    R.inv = {1.0f / d.x, 1.0f / d.y, 1.0f / d.z};
    R.th = 0.0f; R.tu = 0.0f; R.tv = 0.0f;
    R.root_id = root_id;
    R.nxt = 0; R.count = 0;
    R.visits = 0;
    return R;
}

__device__ __forceinline__ pm_hit ray_record(const RayState &R, int hit) {
    if (hit == -1) return miss_record();
    pm_hit h;
    h.id = hit; h.t = R.th; h.u = R.tu; h.v = R.tv;
    return h;
}

// One pass of the tracers' outer loop: start at the root (kd-restart, or
// short stack with an empty stack) or at the popped far child, descend to
// the leaf the ray is in at t = th, move th to the leaf's exit, test the
// leaf.
//   kd-restart   Raytracer.cu:57-106 (TraceNode), :165-262 (KDRestart)
//   short stack  ShortStack.cu:58-194, ShortStack.h:43-98: a ring of the S
//                most recent far children in LDS (interleaved by lane),
//                restart from the root when it runs empty
// descend() is the inner-node half: it returns false when the ray is done
// (a miss) before it reaches a leaf; else node/nd are the leaf and R.th its
// exit.  trace_step() is the whole pass: it returns true when the ray is
// done, `hit` being its triangle or -1.
// The ring lives in one LDS array int[2 * SMAX][TB]: rows 0..SMAX-1 hold the
// far nodes, rows SMAX.. their tNext bits; `stk` points at this lane's column
// and R.nxt counts in rows of TB, so a slot's two words are one register
// offset apart by a constant.
// The descent is split in two so that the cooperative kernel can spread it
// over several loop iterations: pass_start picks the pass's first node (the
// root or the popped far child; false: the visit bound is spent, a miss),
// descend_steps walks at most `cap` inner nodes from `node` (1: a leaf, nd
// its record and R.th its exit; 0: still inner, node/tNext/steps carry on;
// -1: the descent bound is spent, a miss).
template <int SMAX, bool SHORT>
__device__ __forceinline__ bool pass_start(const TreeDev &T, RayState &R, int S, int *stk, int &node,
                                           float &tNext) {
    if (R.visits++ > T.num_nodes) return false;
    if (!SHORT || R.count == 0) {
        node = R.root_id; tNext = F_INF;
    } else {                                     // Pop, ShortStack.h:68-72
        R.nxt = (R.nxt == 0 ? S * TB : R.nxt) - TB;
        R.count--;
        node = stk[R.nxt];
        tNext = __int_as_float(stk[R.nxt + SMAX * TB]);
    }
    return true;
}

template <int SMAX, bool SHORT, bool COUNT>
__device__ __forceinline__ int descend_steps(const TreeDev &T, RayState &R, int S, int *stk, int4 &cnt, int &node,
                                             float &tNext, int &steps, int cap, int4 &nd) {
    nd = T.packed[(unsigned)node];
    for (int k = 0; nd.w != LEAF; ++k) {
        if (k == cap) return 0;
        if (++steps > MAX_DESCENT) return -1;
        if (COUNT) cnt.x++;
        const float s = __int_as_float(nd.x);
        const float ori = comp3(R.o, nd.w), dd = comp3(R.inv, nd.w);
        const float tSplit = (s - ori) * dd;
        const int lower = 0 < dd ? nd.y : nd.z;
        const int upper = 0 < dd ? nd.z : nd.y;
        if (R.th < tSplit) {
            node = lower;
            // here tSplit > th >= 0, and tNext (the sentinel or an earlier
            // tSplit) is never NaN: the select is exactly fminf(tSplit, tNext)
            const bool nearer = tSplit < tNext;
            if (SHORT && nearer) {                   // Push, ShortStack.h:57-62
                stk[R.nxt] = upper;
                stk[R.nxt + SMAX * TB] = __float_as_int(tNext);
                R.nxt = R.nxt == (S - 1) * TB ? 0 : R.nxt + TB;
                R.count = min(R.count + 1, S);       // count <= S: count == S ? S : count + 1
            }
            tNext = nearer ? tSplit : tNext;
        } else {
            node = upper;
        }
        nd = T.packed[(unsigned)node];   // child indices are in [0, num_nodes)
    }
    R.th = tNext * 1.00001f;
    return 1;
}

template <int SMAX, bool SHORT, bool COUNT>
__device__ __forceinline__ bool descend(const TreeDev &T, RayState &R, int S, int *stk, int4 &cnt, int &node,
                                        int4 &nd) {
    float tNext;
    int steps = 0;
    if (!pass_start<SMAX, SHORT>(T, R, S, stk, node, tNext)) return false;
    return descend_steps<SMAX, SHORT, COUNT>(T, R, S, stk, cnt, node, tNext, steps, INT_MAX, nd) > 0;
}

template <int SMAX, bool SHORT, bool SKIP, bool COUNT, bool WOOP>
__device__ __forceinline__ bool trace_step(const TreeDev &T, RayState &R, int S, int *stk, int4 &cnt,
                                           int &hit) {
    hit = -1;
    int node;
    int4 nd;
    if (!descend<SMAX, SHORT, COUNT>(T, R, S, stk, cnt, node, nd)) return true;
    hit = test_leaf<SKIP, COUNT, WOOP>(T, nd, node, R.o, R.d, R.inv, R.th, R.tu, R.tv, cnt);
    return hit != -1 || !(R.th < F_INF);
}

// the whole traversal of one ray (the canvas shading's per-bounce trace)
template <bool SKIP, bool COUNT, bool WOOP>
__device__ __forceinline__ pm_hit kd_restart_ray(const TreeDev &T, int root_id, f3 o, f3 d, int4 &cnt) {
    RayState R = ray_start(T, root_id, o, d);
    int hit;
    while (!trace_step<1, false, SKIP, COUNT, WOOP>(T, R, 1, nullptr, cnt, hit)) {
    }
    return ray_record(R, hit);
}

template <int SMAX, bool SKIP, bool COUNT, bool WOOP>
__device__ __forceinline__ pm_hit short_stack_ray(const TreeDev &T, int root_id, f3 o, f3 d, int S,
                                                  int *stk, int4 &cnt) {
    RayState R = ray_start(T, root_id, o, d);
    int hit;
    while (!trace_step<SMAX, true, SKIP, COUNT, WOOP>(T, R, S, stk, cnt, hit)) {
    }
    return ray_record(R, hit);
}

// The first-hit kernels (kd-restart and short stack): persistent waves.
// A wave takes CHUNK consecutive rays at a time from a global counter
// (*work, zero at launch); once REFILL_MIN of its lanes are done, they take
// the next rays of the wave's chunk, so the wave's lanes stay busy while its
// longest rays finish.  Lane state: the ray's index q and its traversal
// state; one loop iteration = one pass of trace_step for every busy lane.
// The rays are handed out in the order `order` lists them (input order
// without one); ray q reads org[q], dir[q], ray_root[q] and writes out[q].
// Every wave leaves the loop once the counter has
// passed n and its lanes are idle; every ray ends (trace_step's visit and
// descent bounds), so every wave does.
constexpr int CHUNK = 64;        // rays a wave takes from the counter at a time
constexpr int REFILL_MIN = 16;   // idle lanes that make a wave take new rays

__device__ __forceinline__ bool ray_root_valid(const TreeDev &T, const int *ray_root, int r, int &root_id) {
    root_id = ray_root ? ray_root[r] : 0;
    // no such mesh, or (captured batch) the build before this trace failed: a miss
    return !(root_id < 0 || root_id >= T.num_roots || (T.fail && *T.fail));
}

template <int SMAX, bool SHORT, bool SKIP, bool COUNT, bool WOOP>
__global__ __launch_bounds__(TB) void trace_kernel(TreeDev T, const float4 *__restrict__ org,
                                                   const float4 *__restrict__ dir,
                                                   const int *__restrict__ ray_root,
                                                   const unsigned *__restrict__ order, int n, int S,
                                                   pm_hit *__restrict__ out, int4 *__restrict__ counters,
                                                   int *__restrict__ work) {
    __shared__ int stk_rows[2 * (SHORT ? SMAX : 1)][TB];   // short stack: node rows, then tNext rows
    int *const stk = &stk_rows[0][threadIdx.x];
    const int lane = threadIdx.x & 63;
    const unsigned long long below = lane ? (~0ull >> (64 - lane)) : 0ull;
    int next = 0, end = 0;          // the wave's chunk (same in every lane)
    bool exhausted = false;
    int q = -1;                     // this lane's ray, -1 = idle
    RayState R;
    int4 cnt = make_int4(0, 0, 0, 0);
    for (;;) {
        const unsigned long long idle = __ballot(q < 0);
        // refill once a quarter of the lanes are idle: new rays start at
        // the root while the others are deep in the tree, so refilling
        // single lanes costs more than it gains
        if (__popcll(idle) >= REFILL_MIN && !exhausted) {
            if (next >= end) {
                int base = 0;
                if (lane == 0) base = atomicAdd(work, CHUNK);
                base = __builtin_amdgcn_readfirstlane(__shfl(base, 0, 64));
                if (base >= n || base < 0) {
                    exhausted = true;
                } else {
                    next = base;
                    end = n - base < CHUNK ? n : base + CHUNK;
                }
            }
            if (!exhausted) {
                const int avail = end - next;
                const int rank = __popcll(idle & below);
                if (q < 0 && rank < avail) {
                    q = order ? (int)order[next + rank] : next + rank;
                    int root_id;
                    if ((unsigned)q >= (unsigned)n) {
                        q = -1;                  // not a ray index (never for a valid order)
                    } else if (ray_root_valid(T, ray_root, q, root_id)) {
                        R = ray_start(T, root_id, mk3(org[q]), adjust_dir(mk3(dir[q])));
                        cnt = make_int4(0, 0, 0, 0);
                    } else {
                        out[q] = miss_record();
                        if (COUNT) counters[q] = make_int4(0, 0, 0, 0);
                        q = -1;
                    }
                }
                const int took = __popcll(idle);
                next = __builtin_amdgcn_readfirstlane(next + (took < avail ? took : avail));
            }
        }
        const bool busy = q >= 0;
        if (__ballot(busy) == 0) {
            if (exhausted) break;
            continue;
        }
        if (busy) {
            int hit;
            if (trace_step<SMAX, SHORT, SKIP, COUNT, WOOP>(T, R, S, stk, cnt, hit)) {
                out[q] = ray_record(R, hit);
                if (COUNT) counters[q] = cnt;
                q = -1;
            }
        }
    }
}

// ---- cooperative leaf test (Moller-Trumbore tracers)
//
// After a wave's lanes have descended to their leaves, lane i holds a leaf
// with k_i triangles; tested lane by lane the wave would loop max k_i times
// with most lanes idle.  Instead the wave lists all (lane, triangle) pairs
// in LDS (COOP_WIN at a time) and every lane tests the next pair, with the
// owner's ray and leaf exit th (read from the owner, see CoopLds).  An
// owner's first hit is the pair with the smallest t, and among equal t the
// triangle latest in bitmap order -- exactly what the sequential leaf loop
// (accept 0 <= t <= th, then th = t) ends with -- so each accepted pair
// offers the key (t bits, ~slot) to its owner's 64-bit LDS minimum, and the
// owner recomputes its winning triangle itself for t, u, v and the id.
// The key's low word is ~(byte offset of the slot's record): slots of one
// leaf grow with the bit index, so it orders ties like the bitmap does; t >=
// 0 here (-0 keyed as +0).  Record offsets are 32-bit (48 bytes per slot):
// trees with 2^32 / 48 slots or more take the lane-per-ray kernel.
constexpr int COOP_WIN = 512;
static_assert(COOP_WIN % 64 == 0, "a wave reads pair[r + lane] for r < COOP_WIN, r a multiple of 64");

// The tester reads the owner's ray, leaf exit and first slot either from
// LDS records the owners write (RECS) or straight from the owner's registers
// (ds_bpermute, whole wave active).  The short-stack kernel takes the second:
// its LDS, not its registers, caps its occupancy, and the records' 2 KiB a
// wave is what brings a block under a sixth of the CU's LDS.
template <bool RECS>
struct CoopLds {
    alignas(8) unsigned short pair[COOP_WIN];   // bit << 8 | 4 * owner lane
    unsigned long long key[64];
    float4 org[RECS ? 64 : 1];           // owner's origin, w = its leaf's first record offset (bits)
    float4 dir[RECS ? 64 : 1];           // owner's direction, w = its leaf exit th
};

// inclusive prefix sum over the wave's 64 lanes (every lane active): DPP
// row shifts within each row of 16 lanes, then row broadcasts of lanes 15
// and 31 into the rows after them (disabled lanes add the old value, 0)
__device__ __forceinline__ int wave_incl_scan(int v) {
    v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false);   // row_shr:1
    v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false);   // row_shr:2
    v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false);   // row_shr:4
    v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false);   // row_shr:8
    v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false);   // row_bcast:15, rows 1 and 3
    v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false);   // row_bcast:31, rows 2 and 3
    return v;
}

// lane src4 / 4's value of x (every lane of the wave active)
__device__ __forceinline__ float lane_read(float x, int src4) {
    return __int_as_float(__builtin_amdgcn_ds_bpermute(src4, __float_as_int(x)));
}

// orders this lane's LDS accesses with the other lanes' of the wave (a
// wave's LDS instructions execute in order; this keeps the compiler from
// moving them across)
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One (lane, triangle) pair: the owner's ray against the slot whose record
// starts `off` bytes into the leaf triangles; an accepted hit offers its key
// to the owner.  mt_edges' values without its early exits (the lanes test
// different triangles, so a wave never skips a part) and the same
// acceptance, written as min(hx, hy, hz) >= 0, hx <= th, hy + hz <= 1 (fminf
// returns the other operand of a NaN; a NaN hx fails hx <= th, a NaN hy or
// hz fails hy + hz <= 1).
__device__ __forceinline__ void coop_pair(const float4 *__restrict__ L, unsigned long long *key, unsigned off,
                                          f3 ro, f3 rd, float rth) {
    const float4 *R3 = reinterpret_cast<const float4 *>(reinterpret_cast<const char *>(L) + off);
    const float4 a0 = R3[0], a1 = R3[1], a2 = R3[2];
    const f3 v0 = mk3(a0), e1 = mk3(a1), e2 = mk3(a2);
    const f3 p = cross3(rd, e2);
    const f3 t = sub3(ro, v0);
    const f3 q = cross3(t, e1);
    const float inv = rcp_rn(dot3(p, e1));
    const float hx = inv * dot3(q, e2);
    const float hy = inv * dot3(p, t);
    const float hz = inv * dot3(q, rd);
    if (fminf(fminf(hx, hy), hz) >= 0.0f && hx <= rth && hy + hz <= 1.0f)
        atomicMin(key, ((unsigned long long)__float_as_uint(hx + 0.0f) << 32) | (unsigned)~off);
}

// every lane of the wave calls it with the byte offset of its leaf's first
// record (3 float4 per slot; every offset fits 32 bits); returns the byte
// offset of the lane's winning record, or -1
template <bool RECS>
__device__ __forceinline__ long long coop_leaf(const float4 *__restrict__ L, CoopLds<RECS> &C,
                                               unsigned long long bm, unsigned base48, f3 o, f3 d, float th) {
    const int lane = threadIdx.x & 63;
    C.key[lane] = ~0ull;
    if (RECS) {
        C.org[lane] = make_float4(o.x, o.y, o.z, __uint_as_float(base48));
        C.dir[lane] = make_float4(d.x, d.y, d.z, th);
    }
    // k <= 64 pairs per lane: the wave's prefix sum of k (this lane's first
    // pair) and its total (the wave's pairs)
    const int k = __popcll(bm);
    const int incl = wave_incl_scan(k);
    const int excl = incl - k;
    const int W = __builtin_amdgcn_readlane(incl, 63);
    int pos = excl;
    // a pair is bit << 8 | 4 * owner lane (the owner's ds_bpermute address);
    // a run of bits b0 .. b0+k-1 (every leaf the upper stage makes) is listed
    // four pairs to a 64-bit store; other bitmaps bit by bit
    const int lane4 = lane << 2;
    const int b0 = bm ? __builtin_ctzll(bm) : 0;
    const unsigned long long run = bm >> b0;
    const bool contig = (run & (run + 1)) == 0ull;
    unsigned lo = contig ? 0u : (unsigned)bm, hi = contig ? 0u : (unsigned)(bm >> 32);
    unsigned long long *pair4 = reinterpret_cast<unsigned long long *>(C.pair);
    for (int wbase = 0; wbase < W; wbase += COOP_WIN) {
        const int wn = W - wbase < COOP_WIN ? W - wbase : COOP_WIN;
        wave_sync();
        // list this lane's pairs of the window, bits in ascending order
        if (contig) {
            int i = wbase - excl > 0 ? wbase - excl : 0;
            const int iend = wbase + wn - excl < k ? wbase + wn - excl : k;
            int p = excl + i - wbase;
            while (i < iend && (p & 3)) {
                C.pair[p++] = (unsigned short)((b0 + i) << 8 | lane4);
                ++i;
            }
            unsigned long long v = (unsigned long long)((b0 + i) << 8 | lane4) * 0x0001000100010001ull +
                                   0x0300020001000000ull;
            while (i + 4 <= iend) {
                pair4[p >> 2] = v;
                v += 1024ull * 0x0001000100010001ull;
                p += 4;
                i += 4;
            }
            while (i < iend) {
                C.pair[p++] = (unsigned short)((b0 + i) << 8 | lane4);
                ++i;
            }
        }
        while (lo && pos < wbase + wn) {
            const int b = __builtin_ctz(lo);
            lo &= lo - 1;
            C.pair[pos++ - wbase] = (unsigned short)(b << 8 | lane4);
        }
        while (hi && pos < wbase + wn) {
            const int b = __builtin_ctz(hi) + 32;
            hi &= hi - 1;
            C.pair[pos++ - wbase] = (unsigned short)(b << 8 | lane4);
        }
        wave_sync();
        // the owner's ray, leaf exit and first record offset, for pair pr
        auto pair_at = [&](int e, f3 &ro, f3 &rd, float &rth, unsigned &off, int &ow4) {
            const int pr = C.pair[e];   // e < COOP_WIN: stale past wn, used only for an in-range lane
            ow4 = pr & 255;
            unsigned rbase;
            if (RECS) {
                const float4 a = C.org[ow4 >> 2], b = C.dir[ow4 >> 2];
                ro = mk3(a); rd = mk3(b); rth = b.w; rbase = __float_as_uint(a.w);
            } else {
                ro = {lane_read(o.x, ow4), lane_read(o.y, ow4), lane_read(o.z, ow4)};
                rd = {lane_read(d.x, ow4), lane_read(d.y, ow4), lane_read(d.z, ow4)};
                rth = lane_read(th, ow4);
                rbase = (unsigned)__builtin_amdgcn_ds_bpermute(ow4, (int)base48);
            }
            off = (unsigned)(pr >> 8) * 48u + rbase;
        };
        // full rounds of 64 pairs, then the window's tail
        const int full = wn & ~63;
        int r = 0;
        for (; r < full; r += 64) {
            f3 ro, rd;
            float rth;
            unsigned off;
            int ow4;
            pair_at(r + lane, ro, rd, rth, off, ow4);
            coop_pair(L, &C.key[ow4 >> 2], off, ro, rd, rth);
        }
        if (r < wn) {
            f3 ro, rd;
            float rth;
            unsigned off;
            int ow4;
            pair_at(r + lane, ro, rd, rth, off, ow4);
            if (r + lane < wn) coop_pair(L, &C.key[ow4 >> 2], off, ro, rd, rth);
        }
    }
    wave_sync();
    const unsigned long long best = C.key[lane];
    return best == ~0ull ? -1ll : (long long)(unsigned)~(unsigned)(best & 0xffffffffu);
}

// the persistent kernel of trace_kernel with the leaf test done by the wave.
// CAPPED: a lane's descent walks at most `cap` inner nodes per iteration and
// goes on in the next (pm_trace_params.step_cap); the nodes visited, and so
// records and counters, are the same.
template <int SMAX, bool SHORT, bool SKIP, bool COUNT, bool CAPPED>
__global__ __launch_bounds__(TB) void trace_coop_kernel(TreeDev T, const float4 *__restrict__ org,
                                                        const float4 *__restrict__ dir,
                                                        const int *__restrict__ ray_root,
                                                        const unsigned *__restrict__ order, int n, int S,
                                                        pm_hit *__restrict__ out,
                                                        int4 *__restrict__ counters, int *__restrict__ work,
                                                        int cap) {
    __shared__ int stk_rows[2 * (SHORT ? SMAX : 1)][TB];   // short stack: node rows, then tNext rows
    int *const stk = &stk_rows[0][threadIdx.x];
    __shared__ CoopLds<!SHORT> coop[TB / 64];
    const int lane = threadIdx.x & 63;
    CoopLds<!SHORT> &C = coop[threadIdx.x >> 6];
    const unsigned long long below = lane ? (~0ull >> (64 - lane)) : 0ull;
    int next = 0, end = 0;          // the wave's chunk (same in every lane)
    bool exhausted = false;
    int q = -1;                     // this lane's ray, -1 = idle
    // a descent in progress (mid): its node, the far limit tNext, its steps
    bool mid = false;
    int cur = 0, steps = 0;
    float tnext = 0.0f;
    RayState R;
    R.o = {0.0f, 0.0f, 0.0f}; R.d = R.o; R.th = 0.0f;
    int4 cnt = make_int4(0, 0, 0, 0);
    for (;;) {
        const unsigned long long idle = __ballot(q < 0);
        if (__popcll(idle) >= REFILL_MIN && !exhausted) {
            if (next >= end) {
                int b = 0;
                if (lane == 0) b = atomicAdd(work, CHUNK);
                b = __builtin_amdgcn_readfirstlane(__shfl(b, 0, 64));
                if (b >= n || b < 0) {
                    exhausted = true;
                } else {
                    next = b;
                    end = n - b < CHUNK ? n : b + CHUNK;
                }
            }
            if (!exhausted) {
                const int avail = end - next;
                const int rank = __popcll(idle & below);
                if (q < 0 && rank < avail) {
                    q = order ? (int)order[next + rank] : next + rank;
                    int root_id;
                    if ((unsigned)q >= (unsigned)n) {
                        q = -1;                  // not a ray index (never for a valid order)
                    } else if (ray_root_valid(T, ray_root, q, root_id)) {
                        R = ray_start(T, root_id, mk3(org[q]), adjust_dir(mk3(dir[q])));
                        cnt = make_int4(0, 0, 0, 0);
                    } else {
                        out[q] = miss_record();
                        if (COUNT) counters[q] = make_int4(0, 0, 0, 0);
                        q = -1;
                    }
                }
                const int took = __popcll(idle);
                next = __builtin_amdgcn_readfirstlane(next + (took < avail ? took : avail));
            }
        }
        if (__ballot(q >= 0) == 0) {
            if (exhausted) break;
            continue;
        }
        // inner nodes, lane by lane (CAPPED: at most `cap` per iteration)
        bool done = false, at_leaf = false;
        unsigned long long bm = 0;
        int base = 0;
        if (q >= 0) {
            int4 nd;
            int r = 0;
            if (!CAPPED) {
                r = descend<SMAX, SHORT, COUNT>(T, R, S, stk, cnt, cur, nd) ? 1 : -1;
            } else {
                if (!mid) {
                    if (pass_start<SMAX, SHORT>(T, R, S, stk, cur, tnext)) {
                        mid = true;
                        steps = 0;
                    } else {
                        r = -1;
                    }
                }
                if (mid) r = descend_steps<SMAX, SHORT, COUNT>(T, R, S, stk, cnt, cur, tnext, steps, cap, nd);
                if (r != 0) mid = false;
            }
            if (r < 0) {
                done = true;
            } else if (r > 0) {
                const int node = cur;
                at_leaf = true;
                if (COUNT) cnt.y++;
                bool enter = true;
                if (SKIP) {
                    if (COUNT) cnt.w++;
                    enter = ray_box(R.o, R.inv, mk3(T.bmin[node]), mk3(T.bmax[node]));
                }
                if (enter) {
                    bm = ((unsigned long long)(unsigned)nd.z << 32) | (unsigned)nd.y;
                    base = nd.x;
                    if (COUNT) cnt.z += __popcll(bm);
                }
            }
        }
        // the leaves, by the whole wave
        const long long win = coop_leaf(T.leaf_tris, C, bm, (unsigned)base * 48u, R.o, R.d, R.th);
        if (q >= 0 && at_leaf) {
            int hit = -1;
            if (win >= 0) {             // the owner's own test of its winner
                const float4 *W = reinterpret_cast<const float4 *>(reinterpret_cast<const char *>(T.leaf_tris) +
                                                                   (unsigned)win);
                const float4 a0 = W[0];
                if (mt_edges(mk3(a0), mk3(W[1]), mk3(W[2]), R.o, R.d, R.th, R.tu, R.tv))
                    hit = __float_as_int(a0.w);
            }
            done = hit != -1 || !(R.th < F_INF);
            if (done) {
                out[q] = ray_record(R, hit);
                if (COUNT) counters[q] = cnt;
                q = -1;
            }
        } else if (q >= 0 && done) {    // done before a leaf: a miss
            out[q] = miss_record();
            if (COUNT) counters[q] = cnt;
            q = -1;
        }
    }
}

// Brute force: every lane tests every triangle in index order; triangles are
// staged through LDS one tile at a time (one tile = TB triangles, 12 KiB).
// WOOP: P0/P1/P2 are the triangles' Woop matrices m0/m1/m2 instead
template <bool COUNT, bool WOOP>
__global__ __launch_bounds__(TB) void brute_kernel(const float4 *__restrict__ P0,
                                                   const float4 *__restrict__ P1,
                                                   const float4 *__restrict__ P2, int ntris,
                                                   const float4 *__restrict__ org,
                                                   const float4 *__restrict__ dir, int n,
                                                   pm_hit *__restrict__ out,
                                                   int4 *__restrict__ counters) {
    __shared__ float4 s0[TB], s1[TB], s2[TB];
    const int r = blockIdx.x * TB + threadIdx.x;
    const bool live = r < n;
    f3 o = {0.0f, 0.0f, 0.0f}, d = {0.0f, 0.0f, 1.0f};
    if (live) { o = mk3(org[r]); d = mk3(dir[r]); }
    float th = F_INF, tu = 0.0f, tv = 0.0f;
    int hit = -1;
    for (int base = 0; base < ntris; base += TB) {
        const int k = base + threadIdx.x;
        __syncthreads();
        if (k < ntris) { s0[threadIdx.x] = P0[k]; s1[threadIdx.x] = P1[k]; s2[threadIdx.x] = P2[k]; }
        __syncthreads();
        const int m = min(TB, ntris - base);
        if (live) {
            for (int j = 0; j < m; ++j) {
                const bool h = WOOP ? woop_test(s0[j], s1[j], s2[j], o, d, th, tu, tv)
                                    : moller_trumbore(mk3(s0[j]), mk3(s1[j]), mk3(s2[j]), o, d, th, tu, tv);
                if (h) hit = base + j;
            }
        }
    }
    if (!live) return;
    pm_hit res = miss_record();
    if (th < F_INF && hit != -1) { res.id = hit; res.t = th; res.u = tu; res.v = tv; }
    out[r] = res;
    if (COUNT) counters[r] = make_int4(0, 0, ntris, 0);
}

// ---- trace view
// n = n_dev ? *n_dev : n_host (a captured build passes its device node count)
__global__ __launch_bounds__(TB) void pack_nodes_kernel(const int8_t *__restrict__ info,
                                                        const float *__restrict__ split,
                                                        const int2 *__restrict__ children,
                                                        const int *__restrict__ prim_index,
                                                        const long long *__restrict__ bitmap, int n_host,
                                                        const int *n_dev, int4 *__restrict__ out) {
    const int n = n_dev ? *n_dev : n_host;
    for (int i = blockIdx.x * TB + threadIdx.x; i < n; i += gridDim.x * TB) {
        const int ax = info[i] & 3;
        if (ax) {
            const int2 c = children[i];
            out[i] = make_int4(__float_as_int(split[i]), c.x, c.y, ax);
        } else {
            const unsigned long long b = (unsigned long long)bitmap[i];
            out[i] = make_int4(prim_index[i], (int)(unsigned)(b & 0xffffffffull), (int)(unsigned)(b >> 32), LEAF);
        }
    }
}

// leaf slot k -> v0 (w = id bits), e1, e2; with `areas`, also the slot's
// triangle area 0.5 |e1 x e2| (CalcSurfaceArea, LowerTriangleMap.h:24-38,
// the same operations tri_area performs)
__device__ __forceinline__ void leaf_slot(const int *__restrict__ prim_indices, int k,
                                          const float4 *__restrict__ P0, const float4 *__restrict__ P1,
                                          const float4 *__restrict__ P2, int ntris, float4 *__restrict__ out,
                                          int *bad, float *__restrict__ areas) {
    const int t = prim_indices[k];
    if (t < 0 || t >= ntris) {
        if (bad) atomicOr(bad, 1);
        const float nan = __int_as_float(0x7fc00000);
        out[3ll * k] = make_float4(nan, nan, nan, __int_as_float(-1));
        out[3ll * k + 1] = make_float4(nan, nan, nan, 0.0f);
        out[3ll * k + 2] = make_float4(nan, nan, nan, 0.0f);
        return;
    }
    const f3 v0 = mk3(P0[t]), v1 = mk3(P1[t]), v2 = mk3(P2[t]);
    const f3 e1 = sub3(v1, v0), e2 = sub3(v2, v0);
    out[3ll * k] = make_float4(v0.x, v0.y, v0.z, __int_as_float(t));
    out[3ll * k + 1] = make_float4(e1.x, e1.y, e1.z, 0.0f);
    out[3ll * k + 2] = make_float4(e2.x, e2.y, e2.z, 0.0f);
    if (areas) {
        const f3 c = cross3(e1, e2);
        areas[k] = 0.5f * sqrtf(dot3(c, c));
    }
}

// m = m_dev ? *m_dev : m_host
__global__ __launch_bounds__(TB) void leaf_tris_kernel(const int *__restrict__ prim_indices, int m_host,
                                                       const int *m_dev, const float4 *__restrict__ P0,
                                                       const float4 *__restrict__ P1,
                                                       const float4 *__restrict__ P2, int ntris,
                                                       float4 *__restrict__ out, int *bad,
                                                       float *__restrict__ areas) {
    const int m = m_dev ? *m_dev : m_host;
    for (int k = blockIdx.x * TB + threadIdx.x; k < m; k += gridDim.x * TB)
        leaf_slot(prim_indices, k, P0, P1, P2, ntris, out, bad, areas);
}

// the same slots from the triangles' records (a build's prim indices are in
// range): one contiguous 48-byte gather per slot
__global__ __launch_bounds__(TB) void leaf_copy_kernel(const int *__restrict__ prim_indices, int m_host,
                                                       const int *m_dev, const float4 *__restrict__ rec,
                                                       float4 *__restrict__ out, float *__restrict__ areas) {
    const int m = m_dev ? *m_dev : m_host;
    for (int k = blockIdx.x * TB + threadIdx.x; k < m; k += gridDim.x * TB) {
        const int t = prim_indices[k];
        const float4 r0 = rec[3ll * t], r1 = rec[3ll * t + 1], r2 = rec[3ll * t + 2];
        out[3ll * k] = r0;
        out[3ll * k + 1] = r1;
        out[3ll * k + 2] = r2;
        if (areas) {
            const f3 c = cross3(mk3(r1), mk3(r2));
            areas[k] = 0.5f * sqrtf(dot3(c, c));
        }
    }
}

// ---- canvas shading: ColorKernels.h:15-85 in the tracers' shading loops
// (Raytracer.cu:165-262, ShortStack.cu:96-194, BruteTracer.cu:54-96); the
// specular power is the correctly rounded float of the double power
struct Light { f3 pos, ambient, diffuse, specular; };

__device__ __forceinline__ f3 scale3(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ f3 add3(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 mul3(f3 a, f3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ float saturate(float x) { return x > 1.0f ? 1.0f : (x > 0.0f ? x : 0.0f); }
__device__ __forceinline__ float clamp01(float x) { return fmaxf(0.0f, fminf(x, 1.0f)); }

// PhongLighting, ColorKernels.h:15-45, shadow = 1 (the reference's shadow
// rays are disabled, Raytracer.cu:219-226)
__device__ __forceinline__ float4 phong(const float4 c, f3 normal, f3 point, f3 origin, const Light &L) {
    const f3 lightDir = normalize3(sub3(L.pos, point));
    const float diffuse = saturate(dot3(lightDir, normal) * 1.0f);
    f3 refl = sub3(scale3(normal, 2.0f * dot3(normal, lightDir)), lightDir);
    refl = normalize3(refl);
    const float stemp = saturate(dot3(normalize3(sub3(origin, point)), refl));
    const float specProp = 1.0f - c.w;
    const float specular = specProp * (float)pow((double)stemp, (double)(128.0f * specProp));
    const f3 base = {c.x, c.y, c.z};
    const f3 light = add3(mul3(base, add3(L.ambient, scale3(L.diffuse, diffuse))),
                          scale3(scale3(scale3(L.specular, specular), 10.0f), 1.0f));
    const float alpha = c.w < specular ? specular : c.w;
    return make_float4(clamp01(light.x), clamp01(light.y), clamp01(light.z), alpha);
}

// Lighting, ColorKernels.h:47-77: shade the hit, turn the ray into its
// mirror reflection from the hit point
__device__ __forceinline__ float4 lighting(const pm_hit &h, f3 &origin, f3 &dir, const float4 *__restrict__ n0,
                                           const float4 *__restrict__ n1, const float4 *__restrict__ n2,
                                           const uchar4 *__restrict__ c0, const Light &L) {
    const f3 point = add3(origin, scale3(dir, h.t));
    const float w = 1.0f - h.u - h.v;
    f3 normal = add3(add3(scale3(mk3(n0[h.id]), w), scale3(mk3(n1[h.id]), h.u)), scale3(mk3(n2[h.id]), h.v));
    normal = normalize3(normal);
    const uchar4 cc = c0[h.id];
    const float4 col = make_float4(cc.x / 255.0f, cc.y / 255.0f, cc.z / 255.0f, cc.w / 255.0f);
    dir = add3(scale3(normal, 2.0f * dot3(normal, scale3(dir, -1.0f))), dir);
    dir = normalize3(dir);
    const float4 out = phong(col, normal, point, origin, L);
    origin = add3(point, scale3(dir, 0.00001f));
    return out;
}

// BlendColor, ColorKernels.h:79-84
__device__ __forceinline__ float4 blend(const float4 o, float4 n) {
    const float a = n.w;
    n.w = 1.0f;
    const float k = (1.0f - o.w) * a;
    return make_float4(o.x + k * n.x, o.y + k * n.y, o.z + k * n.z, o.w + k * n.w);
}

// brute-force first hit for one ray (Exhaustive, BruteTracer.cu:54-96),
// straight from global memory
template <bool WOOP>
__device__ __forceinline__ pm_hit brute_ray(const float4 *__restrict__ A, const float4 *__restrict__ B,
                                            const float4 *__restrict__ C, int ntris, f3 o, f3 d) {
    float th = F_INF, tu = 0.0f, tv = 0.0f;
    int hit = -1;
    for (int k = 0; k < ntris; ++k) {
        const bool h = WOOP ? woop_test(A[k], B[k], C[k], o, d, th, tu, tv)
                            : moller_trumbore(mk3(A[k]), mk3(B[k]), mk3(C[k]), o, d, th, tu, tv);
        if (h) hit = k;
    }
    pm_hit res = miss_record();
    if (th < F_INF && hit != -1) { res.id = hit; res.t = th; res.u = tu; res.v = tv; }
    return res;
}

// IRayTracer::Trace(canvas): per ray, first hit -> Lighting -> BlendColor
// until a miss, alpha >= 0.97 or max_bounces; one uchar4 per ray
template <int ALGO, bool SKIP, bool WOOP>
__global__ __launch_bounds__(TB) void canvas_kernel(TreeDev T, const float4 *__restrict__ A,
                                                    const float4 *__restrict__ B,
                                                    const float4 *__restrict__ C, int ntris,
                                                    const float4 *__restrict__ org,
                                                    const float4 *__restrict__ dir, int n, int S,
                                                    const float4 *__restrict__ n0,
                                                    const float4 *__restrict__ n1,
                                                    const float4 *__restrict__ n2,
                                                    const uchar4 *__restrict__ c0, Light L, int max_bounces,
                                                    uchar4 *__restrict__ canvas) {
    __shared__ int stk_rows[2 * (ALGO == PM_TRACE_SHORT_STACK ? 16 : 1)][TB];   // node rows, then tNext rows
    int *const stk = &stk_rows[0][threadIdx.x];
    const int r = blockIdx.x * TB + threadIdx.x;
    if (r >= n) return;
    f3 o = mk3(org[r]), d = mk3(dir[r]);
    if (ALGO != PM_TRACE_BRUTE) d = adjust_dir(d);
    float4 col = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    int4 cnt = make_int4(0, 0, 0, 0);
    for (int b = 0; b < max_bounces; ++b) {
        pm_hit h;
        if constexpr (ALGO == PM_TRACE_BRUTE) {
            h = brute_ray<WOOP>(A, B, C, ntris, o, d);
        } else if constexpr (ALGO == PM_TRACE_KD_RESTART) {
            h = kd_restart_ray<SKIP, false, WOOP>(T, 0, o, d, cnt);
        } else {
            h = short_stack_ray<16, SKIP, false, WOOP>(T, 0, o, d, S, stk, cnt);
        }
        if (h.id < 0) break;
        col = blend(col, lighting(h, o, d, n0, n1, n2, c0, L));
        if (ALGO != PM_TRACE_BRUTE) d = adjust_dir(d);
        if (!(col.w < 0.97f)) break;
    }
    canvas[r] = make_uchar4((unsigned char)(col.x * 255.0f), (unsigned char)(col.y * 255.0f),
                            (unsigned char)(col.z * 255.0f), (unsigned char)(col.w * 255.0f));
}

// CreateWoopValues (GeometryList.cu:80-101): matrices of (p1, p2, p0) for
// every triangle (W0/W1/W2 arrays) or for every leaf slot (3 per slot)
__global__ __launch_bounds__(TB) void woop_tris_kernel(const float4 *__restrict__ P0,
                                                       const float4 *__restrict__ P1,
                                                       const float4 *__restrict__ P2, int n,
                                                       float4 *__restrict__ W0, float4 *__restrict__ W1,
                                                       float4 *__restrict__ W2) {
    const int k = blockIdx.x * TB + threadIdx.x;
    if (k >= n) return;
    float4 m0, m1, m2;
    woop_matrix(mk3(P1[k]), mk3(P2[k]), mk3(P0[k]), m0, m1, m2);
    W0[k] = m0; W1[k] = m1; W2[k] = m2;
}

__global__ __launch_bounds__(TB) void woop_slots_kernel(const int *__restrict__ prim_indices, int m,
                                                        const float4 *__restrict__ P0,
                                                        const float4 *__restrict__ P1,
                                                        const float4 *__restrict__ P2, int ntris,
                                                        float4 *__restrict__ W) {
    const int k = blockIdx.x * TB + threadIdx.x;
    if (k >= m) return;
    const int t = prim_indices[k];
    float4 m0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), m1 = m0, m2 = m0;
    if (t >= 0 && t < ntris) woop_matrix(mk3(P1[t]), mk3(P2[t]), mk3(P0[t]), m0, m1, m2);
    W[3ll * k] = m0; W[3ll * k + 1] = m1; W[3ll * k + 2] = m2;
}

__global__ void clear_int_kernel(int *p) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = 0;
}

// the rays of one trace call
struct RayBatch {
    const float4 *org, *dir;
    const int *root;            // per-ray root (forest) or null
    const unsigned *order;      // the order the rays are traced in, or null
    int n;
};

// persistent grid: as many blocks as the device holds at once (fewer for a
// small trace)
struct Launch {
    int num_cus;
    int *work;
    int step_cap;   // pm_trace_params.step_cap (cooperative kernels)
};

template <class K>
int resident_blocks(K kernel, const RayBatch &R, const Launch &L) {
    int b = 0;
    PM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, kernel, TB, 0));
    return (int)std::min<long long>(((long long)R.n + TB - 1) / TB, (long long)(b > 0 ? b : 1) * L.num_cus);
}

template <int SMAX, bool SHORT, bool SKIP, bool COUNT, bool WOOP>
void launch_trace(const TreeDev &T, const RayBatch &R, int S, pm_hit *out, int4 *cnt, const Launch &L,
                  hipStream_t s) {
    if (!WOOP && 48ll * T.num_slots < (1ll << 32)) {
        if (L.step_cap > 0) {
            auto k = trace_coop_kernel<SMAX, SHORT, SKIP, COUNT, true>;
            k<<<resident_blocks(k, R, L), TB, 0, s>>>(T, R.org, R.dir, R.root, R.order, R.n, S, out, cnt, L.work,
                                                      L.step_cap);
        } else {
            auto k = trace_coop_kernel<SMAX, SHORT, SKIP, COUNT, false>;
            k<<<resident_blocks(k, R, L), TB, 0, s>>>(T, R.org, R.dir, R.root, R.order, R.n, S, out, cnt, L.work, 0);
        }
    } else {
        auto k = trace_kernel<SMAX, SHORT, SKIP, COUNT, WOOP>;
        k<<<resident_blocks(k, R, L), TB, 0, s>>>(T, R.org, R.dir, R.root, R.order, R.n, S, out, cnt, L.work);
    }
}

template <bool SKIP, bool COUNT, bool WOOP>
void dispatch(const TreeDev &T, int algo, const RayBatch &R, int S, pm_hit *out, int4 *cnt, const Launch &L,
              hipStream_t s) {
    if (algo == PM_TRACE_KD_RESTART) {
        launch_trace<1, false, SKIP, COUNT, WOOP>(T, R, 1, out, cnt, L, s);
    } else {
        if (S <= 4) launch_trace<4, true, SKIP, COUNT, WOOP>(T, R, S, out, cnt, L, s);
        else if (S <= 8) launch_trace<8, true, SKIP, COUNT, WOOP>(T, R, S, out, cnt, L, s);
        else launch_trace<16, true, SKIP, COUNT, WOOP>(T, R, S, out, cnt, L, s);
    }
}

template <bool WOOP>
void dispatch_flags(const TreeDev &T, bool skip, bool count, int algo, const RayBatch &R, int S, pm_hit *out,
                    int4 *cnt, const Launch &L, hipStream_t s) {
    if (skip) {
        if (count) dispatch<true, true, WOOP>(T, algo, R, S, out, cnt, L, s);
        else dispatch<true, false, WOOP>(T, algo, R, S, out, cnt, L, s);
    } else {
        if (count) dispatch<false, true, WOOP>(T, algo, R, S, out, cnt, L, s);
        else dispatch<false, false, WOOP>(T, algo, R, S, out, cnt, L, s);
    }
}

}  // namespace

void build_node_view(pm_ctx *ctx, hipStream_t s, const int *n_dev) {
    const int n = ctx->num_nodes;
    ctx->packed.reserve((size_t)n + 1, s);
    if (n > 0) {
        const int blocks = std::min((n + TB - 1) / TB, n_dev ? 4096 : INT_MAX);
        pack_nodes_kernel<<<blocks, TB, 0, s>>>(ctx->nodes.info.ptr, ctx->nodes.split.ptr, ctx->nodes.children.ptr,
                                                ctx->nodes.prim_index.ptr, ctx->nodes.bitmap.ptr, n, n_dev,
                                                ctx->packed.ptr);
        PM_LAUNCH_CHECK();
    }
}

void build_leaf_view(pm_ctx *ctx, float *areas, bool check, hipStream_t s, const int *m_dev, const float4 *rec) {
    const int m = ctx->num_prim_indices;
    ctx->leaf_tris.reserve(3 * (size_t)m + 3, s);
    if (m == 0) return;
    BuildScratch &sc = ctx->scratch;
    int *bad = nullptr;
    if (check) {
        sc.totals.reserve(8, s);
        sc.host.ensure(8);
        clear_int_kernel<<<1, 64, 0, s>>>(sc.totals.ptr);
        PM_LAUNCH_CHECK();
        bad = sc.totals.ptr;
    }
    const int blocks = std::min((m + TB - 1) / TB, m_dev ? 8192 : INT_MAX);
    if (rec && !check)
        leaf_copy_kernel<<<blocks, TB, 0, s>>>(ctx->prim_indices.ptr, m, m_dev, rec, ctx->leaf_tris.ptr, areas);
    else
        leaf_tris_kernel<<<blocks, TB, 0, s>>>(ctx->prim_indices.ptr, m, m_dev, ctx->p0, ctx->p1, ctx->p2,
                                               ctx->ntris, ctx->leaf_tris.ptr, bad, areas);
    PM_LAUNCH_CHECK();
    if (check) {
        PM_HIP(hipMemcpyAsync(sc.host.ptr, sc.totals.ptr, sizeof(int), hipMemcpyDeviceToHost, s));
        PM_HIP(hipStreamSynchronize(s));
        if (sc.host.ptr[0]) throw Error{PM_ERR_ARG, "prim index outside the triangle array"};
    }
}

void build_trace_view(pm_ctx *ctx, hipStream_t s) {
    build_node_view(ctx, s);
    build_leaf_view(ctx, nullptr, true, s);
}

namespace {

TreeDev tree_dev(pm_ctx *ctx) {
    TreeDev T;
    T.info = ctx->nodes.info.ptr;
    T.split = ctx->nodes.split.ptr;
    T.children = ctx->nodes.children.ptr;
    T.prim_index = ctx->nodes.prim_index.ptr;
    T.bitmap = ctx->nodes.bitmap.ptr;
    T.bmin = ctx->nodes.bmin.ptr;
    T.bmax = ctx->nodes.bmax.ptr;
    T.prim_indices = ctx->prim_indices.ptr;
    T.packed = ctx->packed.ptr;
    T.leaf_tris = ctx->leaf_tris.ptr;
    T.leaf_woop = nullptr;
    T.num_nodes = ctx->num_nodes;
    T.num_roots = ctx->num_roots;
    T.num_slots = ctx->num_prim_indices;
    T.fail = ctx->trace_fail;
    return T;
}

// GetWoopValues, computed per call: per leaf slot for the tree tracers
const float4 *woop_slots(pm_ctx *ctx, hipStream_t s) {
    const int m = ctx->num_prim_indices;
    ctx->woop.reserve(3 * (size_t)m + 3, s);
    if (m > 0) {
        woop_slots_kernel<<<(m + TB - 1) / TB, TB, 0, s>>>(ctx->prim_indices.ptr, m, ctx->p0, ctx->p1, ctx->p2,
                                                           ctx->ntris, ctx->woop.ptr);
        PM_LAUNCH_CHECK();
    }
    return ctx->woop.ptr;
}

// ... and per triangle (W0 | W1 | W2) for brute force
float4 *woop_tris(pm_ctx *ctx, int tris, hipStream_t s) {
    ctx->woop.reserve(3 * (size_t)tris + 3, s);
    float4 *W = ctx->woop.ptr;
    if (tris > 0) {
        woop_tris_kernel<<<(tris + TB - 1) / TB, TB, 0, s>>>(ctx->p0, ctx->p1, ctx->p2, tris, W, W + tris,
                                                             W + 2ll * tris);
        PM_LAUNCH_CHECK();
    }
    return W;
}

template <int ALGO>
void launch_canvas(bool skip, bool woop, int blocks, hipStream_t s, const TreeDev &T, const float4 *A,
                   const float4 *B, const float4 *C, int ntris, const float4 *org, const float4 *dir, int n,
                   int S, const float4 *n0, const float4 *n1, const float4 *n2, const uchar4 *c0,
                   const Light &L, int maxb, uchar4 *canvas) {
    if (skip && woop)
        canvas_kernel<ALGO, true, true><<<blocks, TB, 0, s>>>(T, A, B, C, ntris, org, dir, n, S, n0, n1, n2, c0, L, maxb, canvas);
    else if (skip)
        canvas_kernel<ALGO, true, false><<<blocks, TB, 0, s>>>(T, A, B, C, ntris, org, dir, n, S, n0, n1, n2, c0, L, maxb, canvas);
    else if (woop)
        canvas_kernel<ALGO, false, true><<<blocks, TB, 0, s>>>(T, A, B, C, ntris, org, dir, n, S, n0, n1, n2, c0, L, maxb, canvas);
    else
        canvas_kernel<ALGO, false, false><<<blocks, TB, 0, s>>>(T, A, B, C, ntris, org, dir, n, S, n0, n1, n2, c0, L, maxb, canvas);
}

}  // namespace

void trace_rays(pm_ctx *ctx, int algo, const float4 *org, const float4 *dir, int nrays,
                const int *ray_mesh, pm_hit *out, const pm_trace_params &prm, int4 *counters,
                hipStream_t s, const StaticScratch *ext) {
    if (nrays == 0) return;
    const bool woop = prm.intersection == PM_INTERSECT_WOOP;
    if (algo == PM_TRACE_BRUTE) {
        const int tris = prm.brute_cap > 0 ? std::min(ctx->ntris, prm.brute_cap) : ctx->ntris;
        const int blocks = (nrays + TB - 1) / TB;
        const float4 *A = ctx->p0, *B = ctx->p1, *C = ctx->p2;
        if (woop) {
            const float4 *W = woop_tris(ctx, tris, s);
            A = W; B = W + tris; C = W + 2ll * tris;
            if (counters) brute_kernel<true, true><<<blocks, TB, 0, s>>>(A, B, C, tris, org, dir, nrays, out, counters);
            else brute_kernel<false, true><<<blocks, TB, 0, s>>>(A, B, C, tris, org, dir, nrays, out, counters);
        } else {
            if (counters) brute_kernel<true, false><<<blocks, TB, 0, s>>>(A, B, C, tris, org, dir, nrays, out, counters);
            else brute_kernel<false, false><<<blocks, TB, 0, s>>>(A, B, C, tris, org, dir, nrays, out, counters);
        }
        PM_LAUNCH_CHECK();
        return;
    }
    TreeDev T = tree_dev(ctx);
    if (woop) T.leaf_woop = woop_slots(ctx, s);
    const bool skip = prm.leaf_skipping && ctx->has_boxes;
    const int S = prm.stack_size;
    RayBatch R{org, dir, ray_mesh, nullptr, nrays};
    if (ray_order_applies(prm, nrays))
        R.order = ray_order(ctx, org, dir, ray_mesh, nrays, s, ext ? &ext->order : nullptr);
    if (!ctx->num_cus) {
        int v = 0;
        PM_HIP(hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, ctx->device));
        ctx->num_cus = v > 0 ? v : 1;
    }
    int *work = nullptr;
    if (ext) {
        work = ext->work;
    } else {
        ctx->scratch.trace_work.reserve(1, s);
        work = ctx->scratch.trace_work.ptr;
    }
    clear_int_kernel<<<1, 64, 0, s>>>(work);
    PM_LAUNCH_CHECK();
    const Launch L{ctx->num_cus, work, prm.step_cap};
    if (woop) dispatch_flags<true>(T, skip, counters != nullptr, algo, R, S, out, counters, L, s);
    else dispatch_flags<false>(T, skip, counters != nullptr, algo, R, S, out, counters, L, s);
    PM_LAUNCH_CHECK();
}

void trace_canvas(pm_ctx *ctx, int algo, const float4 *org, const float4 *dir, int nrays, const float4 *n0,
                  const float4 *n1, const float4 *n2, const uchar4 *c0, const pm_light &light, int max_bounces,
                  uchar4 *canvas, const pm_trace_params &prm, hipStream_t s) {
    if (nrays == 0) return;
    const bool woop = prm.intersection == PM_INTERSECT_WOOP;
    const Light L = {{light.position[0], light.position[1], light.position[2]},
                     {light.ambient[0], light.ambient[1], light.ambient[2]},
                     {light.diffuse[0], light.diffuse[1], light.diffuse[2]},
                     {light.specular[0], light.specular[1], light.specular[2]}};
    const int blocks = (nrays + TB - 1) / TB;
    TreeDev T = tree_dev(ctx);
    const float4 *A = ctx->p0, *B = ctx->p1, *C = ctx->p2;
    int tris = 0;
    if (algo == PM_TRACE_BRUTE) {
        tris = prm.brute_cap > 0 ? std::min(ctx->ntris, prm.brute_cap) : ctx->ntris;
        if (woop) {
            const float4 *W = woop_tris(ctx, tris, s);
            A = W; B = W + tris; C = W + 2ll * tris;
        }
        launch_canvas<PM_TRACE_BRUTE>(false, woop, blocks, s, T, A, B, C, tris, org, dir, nrays, 1, n0, n1, n2, c0,
                                      L, max_bounces, canvas);
    } else {
        if (woop) T.leaf_woop = woop_slots(ctx, s);
        const bool skip = prm.leaf_skipping && ctx->has_boxes;
        if (algo == PM_TRACE_KD_RESTART)
            launch_canvas<PM_TRACE_KD_RESTART>(skip, woop, blocks, s, T, A, B, C, 0, org, dir, nrays, 1, n0, n1, n2,
                                               c0, L, max_bounces, canvas);
        else
            launch_canvas<PM_TRACE_SHORT_STACK>(skip, woop, blocks, s, T, A, B, C, 0, org, dir, nrays,
                                                prm.stack_size, n0, n1, n2, c0, L, max_bounces, canvas);
    }
    PM_LAUNCH_CHECK();
}

}  // namespace pmkd
