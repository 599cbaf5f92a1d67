// pm_order.hip -- coherent ray order for the tree tracers, and the stable
// LSD radix sort of (key, value) pairs it shares with the backward pass.
//
// Incoherent rays (photons, diffuse bounces, random probes) leave most lanes
// of a wave idle while one lane walks a long path, and spread a wave's node
// and triangle reads over the whole tree.  Before a large trace the rays are
// put in the order of a 32-bit spatial key -- the Morton code of the origin
// cell (6 bits per axis, relative to the ray's root box) above the Morton
// code of the octahedral direction cell (7 bits per axis) -- and the
// tracers hand the rays to their lanes in that order.  Every ray is still
// traced on its own, from the same inputs, by the same code, and its record
// is written to its own index, so
// hits and counters do not depend on the order (the tracers' records are
// bit-identical either way).  In a forest the root index sits above the
// spatial bits.  Not in the reference (its tracers run rays in pixel order,
// Raytracer.cu:295-390, ShortStack.cu:231-321).
//
// Radix sort: 8-bit digits, least significant first.  Per digit: a
// per-tile histogram (4096 keys per tile; the key kernel makes the first
// digit's), a scan of each digit's row of the digit-major histogram table
// (one block per digit, giving the tiles' offsets inside the digit and the
// digit's total), and a scatter that adds the digit totals' prefix, ranks
// keys inside each wave's 1024 consecutive keys with ballot multisplit,
// stages the tile in LDS in digit order and writes each digit's run
// contiguously.  Stable, so LSD passes compose.
#include <climits>

#include "pm_device.h"
#include "pm_internal.h"

namespace pmkd {

namespace {

constexpr int TB = 256;                          // threads per block (4 waves)
constexpr int ITEMS = 16;                        // keys per thread and tile
constexpr int TILE = TB * ITEMS;                 // keys per sort tile
constexpr int WAVE_KEYS = TILE / (TB / 64);      // consecutive keys ranked by one wave
constexpr int ORDER_MIN_RAYS = 1 << 17;          // below this the input order is kept
constexpr int ORDER_BITS = 32;

__device__ __forceinline__ unsigned spread3(unsigned x) {   // 6 bits -> bits 0, 3, ..., 15
    unsigned r = 0;
#pragma unroll
    for (int i = 0; i < 6; ++i) r |= ((x >> i) & 1u) << (3 * i);
    return r;
}

__device__ __forceinline__ unsigned spread2(unsigned x) {   // 7 bits -> bits 0, 2, ..., 12
    unsigned r = 0;
#pragma unroll
    for (int i = 0; i < 7; ++i) r |= ((x >> i) & 1u) << (2 * i);
    return r;
}

// cell index of x in [0, 1) at `cells` cells; NaN and out-of-range clamp
__device__ __forceinline__ unsigned cell(float x, float cells) {
    const float c = fminf(fmaxf(x * cells, 0.0f), cells - 1.0f);
    return (unsigned)c;
}

__device__ __forceinline__ unsigned spatial_key(float4 o, float4 d, float4 lo, float4 hi) {
    const unsigned ox = cell((o.x - lo.x) / (hi.x - lo.x), 64.0f);
    const unsigned oy = cell((o.y - lo.y) / (hi.y - lo.y), 64.0f);
    const unsigned oz = cell((o.z - lo.z) / (hi.z - lo.z), 64.0f);
    // octahedral map of the direction onto [-1, 1]^2
    const float a = fabsf(d.x) + fabsf(d.y) + fabsf(d.z);
    float u = d.x / a, v = d.y / a;
    if (d.z < 0.0f) {
        const float uu = (1.0f - fabsf(v)) * (u >= 0.0f ? 1.0f : -1.0f);
        v = (1.0f - fabsf(u)) * (v >= 0.0f ? 1.0f : -1.0f);
        u = uu;
    }
    const unsigned du = cell((u + 1.0f) * 0.5f, 128.0f);
    const unsigned dv = cell((v + 1.0f) * 0.5f, 128.0f);
    const unsigned km = spread3(ox) | (spread3(oy) << 1) | (spread3(oz) << 2);   // 18 bits
    const unsigned kd = spread2(du) | (spread2(dv) << 1);                         // 14 bits
    return (km << 14) | kd;
}

// keys of a tile of rays + the histogram of their lowest digit
__global__ __launch_bounds__(TB) void ray_key_kernel(const float4 *__restrict__ bmin,
                                                     const float4 *__restrict__ bmax, int num_roots,
                                                     int root_bits, const int *__restrict__ ray_root,
                                                     const float4 *__restrict__ org,
                                                     const float4 *__restrict__ dir, int n, int nb,
                                                     unsigned *__restrict__ keys, int *__restrict__ hist) {
    __shared__ int cnt[256];
    cnt[threadIdx.x] = 0;
    __syncthreads();
    const long long tile = (long long)blockIdx.x * TILE;
    for (int it = 0; it < ITEMS; ++it) {
        const long long i = tile + it * TB + threadIdx.x;
        if (i >= n) break;
        const int root = ray_root ? ray_root[i] : 0;
        unsigned key = 0xffffffffu;                 // rays of no tree go last
        if (root >= 0 && root < num_roots) {
            const unsigned sp = spatial_key(org[i], dir[i], bmin[root], bmax[root]);
            key = root_bits ? (((unsigned)root << (32 - root_bits)) | (sp >> root_bits)) : sp;
        }
        keys[i] = key;
        atomicAdd(&cnt[key & 255u], 1);
    }
    __syncthreads();
    hist[threadIdx.x * nb + blockIdx.x] = cnt[threadIdx.x];    // digit-major
}

__global__ __launch_bounds__(TB) void radix_hist_kernel(const unsigned *__restrict__ keys, int n, int shift,
                                                        int nb, int *__restrict__ hist) {
    __shared__ int cnt[256];
    cnt[threadIdx.x] = 0;
    __syncthreads();
    const long long tile = (long long)blockIdx.x * TILE;
    for (int it = 0; it < ITEMS; ++it) {
        const long long i = tile + it * TB + threadIdx.x;
        if (i >= n) break;
        atomicAdd(&cnt[(keys[i] >> shift) & 255u], 1);
    }
    __syncthreads();
    hist[threadIdx.x * nb + blockIdx.x] = cnt[threadIdx.x];
}

// block d: exclusive prefix of hist[d][0..nb) into offs (in place allowed)
// and the digit's total into dtot[d]
__global__ __launch_bounds__(TB) void radix_digit_scan_kernel(const int *__restrict__ hist, int nb,
                                                              int *__restrict__ offs, int *__restrict__ dtot) {
    __shared__ int wsum[TB / 64];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const long long row = (long long)blockIdx.x * nb;
    int carry = 0;
    for (int base = 0; base < nb; base += TB * 4) {
        int v[4], sum = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int b = base + 4 * t + q;
            v[q] = b < nb ? hist[row + b] : 0;
            sum += v[q];
        }
        int incl = sum;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int y = __shfl_up(incl, d, 64);
            if (lane >= d) incl += y;
        }
        if (lane == 63) wsum[w] = incl;
        __syncthreads();
        int before = carry, chunk = 0;
        for (int q = 0; q < TB / 64; ++q) {
            if (q < w) before += wsum[q];
            chunk += wsum[q];
        }
        int run = before + incl - sum;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int b = base + 4 * t + q;
            if (b < nb) offs[row + b] = run;
            run += v[q];
        }
        carry += chunk;
        __syncthreads();
    }
    if (t == 0) dtot[blockIdx.x] = carry;
}

// one stable pass: tile b's keys with digit d go to
// (sum of the totals of digits < d) + offs[d * nb + b] onwards, in tile
// order.  IOTA: the values are the keys' input indices.
template <bool IOTA>
__global__ __launch_bounds__(TB) void radix_scatter_kernel(const unsigned *__restrict__ kin,
                                                           const unsigned *__restrict__ vin, int n, int shift,
                                                           int nb, const int *__restrict__ offs,
                                                           const int *__restrict__ dtot,
                                                           unsigned *__restrict__ kout,
                                                           unsigned *__restrict__ vout) {
    __shared__ int wcnt[TB / 64][256];     // per wave: keys of each digit so far, then the wave's start
    __shared__ int dstart[256];            // digit start inside the staged tile
    __shared__ int gbase[256];             // digit start in the output
    __shared__ int wsum[TB / 64];
    __shared__ unsigned skey[TILE], sval[TILE];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    for (int i = t; i < (TB / 64) * 256; i += TB) (&wcnt[0][0])[i] = 0;
    __syncthreads();
    const long long tile = (long long)blockIdx.x * TILE;
    const int cnt = (int)((long long)n - tile < TILE ? (long long)n - tile : TILE);
    const unsigned long long below = lane ? (~0ull >> (64 - lane)) : 0ull;
    volatile int *wc = wcnt[w];
    unsigned k[ITEMS], v[ITEMS];
    int rk[ITEMS];
#pragma unroll
    for (int it = 0; it < ITEMS; ++it) {
        const int j = w * WAVE_KEYS + it * 64 + lane;       // tile order = (wave, item, lane)
        const bool valid = j < cnt;
        unsigned key = 0, val = 0;
        if (valid) {
            key = kin[tile + j];
            val = IOTA ? (unsigned)(tile + j) : vin[tile + j];
        }
        const unsigned dg = (key >> shift) & 255u;
        unsigned long long peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const unsigned long long bb = __ballot((dg >> b) & 1u);
            peers &= ((dg >> b) & 1u) ? bb : ~bb;
        }
        const int r = __popcll(peers & below);
        // every lane reads its digit's count before the group's first lane
        // adds the group (LDS instructions of one wave execute in order)
        const int prior = valid ? wc[dg] : 0;
        if (valid && r == 0) wc[dg] = prior + __popcll(peers);
        k[it] = key; v[it] = val; rk[it] = prior + r;
    }
    __syncthreads();
    // thread t owns digit t: the digit's start in the output (prefix of the
    // digit totals), the waves' starts, the tile's digit total and its scan
    {
        const int dt = dtot[t];
        int incl = dt;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int y = __shfl_up(incl, d, 64);
            if (lane >= d) incl += y;
        }
        if (lane == 63) wsum[w] = incl;
        __syncthreads();
        int before = 0;
        for (int q = 0; q < w; ++q) before += wsum[q];
        gbase[t] = before + incl - dt + offs[t * nb + blockIdx.x];
        __syncthreads();
    }
    int total = 0;
    for (int q = 0; q < TB / 64; ++q) {
        const int c = wcnt[q][t];
        wcnt[q][t] = total;
        total += c;
    }
    int incl = total;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int y = __shfl_up(incl, d, 64);
        if (lane >= d) incl += y;
    }
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    int before = 0;
    for (int q = 0; q < w; ++q) before += wsum[q];
    const int ex = before + incl - total;
    dstart[t] = ex;
    __syncthreads();
#pragma unroll
    for (int it = 0; it < ITEMS; ++it) {
        const int j = w * WAVE_KEYS + it * 64 + lane;
        if (j < cnt) {
            const unsigned dg = (k[it] >> shift) & 255u;
            const int pos = dstart[dg] + wcnt[w][dg] + rk[it];
            skey[pos] = k[it];
            sval[pos] = v[it];
        }
    }
    __syncthreads();
    for (int i = t; i < cnt; i += TB) {
        const unsigned key = skey[i];
        const unsigned dg = (key >> shift) & 255u;
        const int g = gbase[dg] + (i - dstart[dg]);
        kout[g] = key;
        vout[g] = sval[i];
    }
}

int sort_tiles(long long n) { return (int)((n + TILE - 1) / TILE); }

}  // namespace

int radix_sort_pairs(unsigned *const keys[2], unsigned *const vals[2], int n, int bits, int *hist, int *offs,
                     bool first_hist_ready, bool iota, BuildScratch &sc, hipStream_t s) {
    if (n <= 0) return 0;
    const int nb = sort_tiles(n);
    int *dtot = offs + 256 * (size_t)nb;        // 256 digit totals after the offsets
    int cur = 0;
    for (int shift = 0; shift < bits; shift += 8) {
        if (shift > 0 || !first_hist_ready) {
            radix_hist_kernel<<<nb, TB, 0, s>>>(keys[cur], n, shift, nb, hist);
            PM_LAUNCH_CHECK();
        }
        radix_digit_scan_kernel<<<256, TB, 0, s>>>(hist, nb, offs, dtot);
        PM_LAUNCH_CHECK();
        if (iota && shift == 0)
            radix_scatter_kernel<true><<<nb, TB, 0, s>>>(keys[cur], vals[cur], n, shift, nb, offs, dtot,
                                                         keys[cur ^ 1], vals[cur ^ 1]);
        else
            radix_scatter_kernel<false><<<nb, TB, 0, s>>>(keys[cur], vals[cur], n, shift, nb, offs, dtot,
                                                          keys[cur ^ 1], vals[cur ^ 1]);
        PM_LAUNCH_CHECK();
        cur ^= 1;
    }
    return cur;
}

size_t radix_hist_size(long long n) { return 256 * (size_t)sort_tiles(n) + 256; }

bool ray_order_applies(const pm_trace_params &prm, int nrays) {
    return prm.ray_order == PM_RAY_ORDER_COHERENT && nrays >= ORDER_MIN_RAYS;
}

void reserve_ray_order(pm_ctx *ctx, int nrays, hipStream_t s) {
    BuildScratch &sc = ctx->scratch;
    for (int i = 0; i < 2; ++i) {
        sc.order_keys[i].reserve((size_t)nrays + 1, s);
        sc.order_vals[i].reserve((size_t)nrays + 1, s);
    }
    sc.order_hist.reserve(radix_hist_size(nrays), s);
    sc.order_offs.reserve(radix_hist_size(nrays), s);
}

const unsigned *ray_order(pm_ctx *ctx, const float4 *org, const float4 *dir, const int *ray_root, int nrays,
                          hipStream_t s, const OrderScratch *ext) {
    BuildScratch &sc = ctx->scratch;
    OrderScratch o;
    if (ext) {
        o = *ext;
    } else {
        reserve_ray_order(ctx, nrays, s);
        o = {{sc.order_keys[0].ptr, sc.order_keys[1].ptr}, {sc.order_vals[0].ptr, sc.order_vals[1].ptr},
             sc.order_hist.ptr, sc.order_offs.ptr};
    }
    const int nb = sort_tiles(nrays);
    int root_bits = 0;
    while (root_bits < 31 && ((unsigned)(ctx->num_roots - 1) >> root_bits)) ++root_bits;
    ray_key_kernel<<<nb, TB, 0, s>>>(ctx->nodes.bmin.ptr, ctx->nodes.bmax.ptr, ctx->num_roots, root_bits,
                                     ray_root, org, dir, nrays, nb, o.keys[0], o.hist);
    PM_LAUNCH_CHECK();
    unsigned *const K[2] = {o.keys[0], o.keys[1]};
    unsigned *const V[2] = {o.vals[0], o.vals[1]};
    return V[radix_sort_pairs(K, V, nrays, ORDER_BITS, o.hist, o.offs, true, true, sc, s)];
}

}  // namespace pmkd
