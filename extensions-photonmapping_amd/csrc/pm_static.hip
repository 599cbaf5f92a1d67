// pm_static.hip -- the scratch of the allocation-free trace (pm_trace_noalloc).
//
// Callers that cannot allocate device memory (code under HIP graph capture,
// tight serving loops) trace through pm_trace_noalloc, whose kernels keep
// their scratch -- the tracers' work counter and, for the coherent ray
// order, the radix sort's keys, values and histograms -- in the file-scope
// device array below instead of context buffers.  The array lives in the
// code object, so every GPU the library runs on holds its own copy until the
// process ends (about 17 MB); its address on a device is looked up once, with
// that device current, and kept per device.  In host code `scratch` names a
// host-side stand-in, never the device memory: only the looked-up pointers
// are passed to kernels.  Every call on a GPU shares that GPU's array:
// calls that overlap on one GPU (two streams, two threads) race.
#include <mutex>

#include "pm_internal.h"

namespace pmkd {

namespace {

constexpr size_t ALIGN = 256;
constexpr size_t align_up(size_t x) { return (x + ALIGN - 1) / ALIGN * ALIGN; }

// layout: work counter | keys[0] | keys[1] | vals[0] | vals[1] | hist | offs
constexpr size_t N_ORDER = (size_t)NOALLOC_MAX_ORDERED_RAYS + 1;
constexpr size_t TILES = ((size_t)NOALLOC_MAX_ORDERED_RAYS + 4095) / 4096;   // sort tiles of 4096 keys
constexpr size_t HIST = 256 * TILES + 256;                                   // radix_hist_size
constexpr size_t OFF_WORK = 0;
constexpr size_t OFF_KEYS = align_up(sizeof(int));
constexpr size_t ARR = align_up(N_ORDER * sizeof(unsigned));
constexpr size_t OFF_HIST = OFF_KEYS + 4 * ARR;
constexpr size_t OFF_OFFS = OFF_HIST + align_up(HIST * sizeof(int));
constexpr size_t SCRATCH_BYTES = OFF_OFFS + align_up(HIST * sizeof(int));

static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_lock;
StaticScratch g_dev[MAX_DEVICES];
bool g_ready[MAX_DEVICES] = {};

}  // namespace

const StaticScratch &static_scratch(int device) {
    if (device < 0 || device >= MAX_DEVICES) throw Error{PM_ERR_ARG, "device index out of range"};
    std::lock_guard<std::mutex> lk(g_lock);
    if (!g_ready[device]) {
        unsigned char *base = nullptr;
        const hipError_t e = hipGetSymbolAddress(reinterpret_cast<void **>(&base), HIP_SYMBOL(scratch));
        if (e != hipSuccess || !base)
            throw Error{PM_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e)};
        if (radix_hist_size(NOALLOC_MAX_ORDERED_RAYS) != HIST)
            throw Error{PM_ERR_INTERNAL, "static scratch histogram size mismatch"};
        StaticScratch &d = g_dev[device];
        d.work = reinterpret_cast<int *>(base + OFF_WORK);
        for (int k = 0; k < 2; ++k) {
            d.order.keys[k] = reinterpret_cast<unsigned *>(base + OFF_KEYS + k * ARR);
            d.order.vals[k] = reinterpret_cast<unsigned *>(base + OFF_KEYS + (2 + k) * ARR);
        }
        d.order.hist = reinterpret_cast<int *>(base + OFF_HIST);
        d.order.offs = reinterpret_cast<int *>(base + OFF_OFFS);
        g_ready[device] = true;
    }
    return g_dev[device];
}

}  // namespace pmkd
