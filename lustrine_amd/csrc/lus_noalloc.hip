// lus_noalloc.hip -- the stable counting sort on caller-owned device buffers,
// with every intermediate in one static device array: the entry point for
// callers that cannot allocate device memory (code under HIP graph capture,
// serving loops, harnesses that forbid hipMalloc).
//
// `scratch` below is a file-scope __device__ array.  In host code its name is
// a host-side shadow, not a device address: the device address is looked up
// with hipGetSymbolAddress once per device, with that device current, and only
// that pointer is passed to kernels.  Each GPU holds its own copy of the
// array, which keeps what the last call on that GPU left in it: the histogram
// is cleared by a fill kernel at the start of every call, and calls that
// overlap on one GPU race on it.
#include <hip/hip_runtime.h>

#include <mutex>

#include "lus_device.hpp"
#include "lus_internal.hpp"
#include "lus_launch.hpp"
#include "lustrine_gpu.h"

namespace {

// scratch layout (ints, each section 64-int aligned):
//   counts [LUS_NOALLOC_MAX_CELLS + 1]   histogram, must start at zero
//   slot   [LUS_NOALLOC_MAX_KEYS]        rank of each key within its cell (arrival order)
//   tmp    [LUS_NOALLOC_MAX_KEYS]        slot -> index before the stable rank
//   ks     [LUS_NOALLOC_MAX_KEYS]        keys in sorted order
//   tiles  [scan tiles of MAX_CELLS + 1, + 1]
constexpr size_t kAlignInts = 64;
constexpr size_t align_up(size_t x) { return (x + kAlignInts - 1) / kAlignInts * kAlignInts; }
constexpr size_t kCountsOff = 0;
constexpr size_t kSlotOff = align_up(kCountsOff + (size_t)LUS_NOALLOC_MAX_CELLS + 1);
constexpr size_t kTmpOff = align_up(kSlotOff + (size_t)LUS_NOALLOC_MAX_KEYS);
constexpr size_t kKsOff = align_up(kTmpOff + (size_t)LUS_NOALLOC_MAX_KEYS);
constexpr size_t kTilesOff = align_up(kKsOff + (size_t)LUS_NOALLOC_MAX_KEYS);
// 2,048-int scan tiles (lus_kernels.hip); a generous bound on their count
constexpr size_t kTilesInts = ((size_t)LUS_NOALLOC_MAX_CELLS + 1) / 512 + 64;
constexpr size_t kScratchInts = align_up(kTilesOff + kTilesInts);

__device__ __align__(256) int scratch[kScratchInts];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
int* g_scratch_ptr[kMaxDevices] = {};

// the device address of this device's `scratch`, looked up on its first use
// here with the device current
int scratch_for_current_device(int* out_dev, int** out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return lus::report_error(LUS_EHIP, "lus_counting_sort_noalloc: hipGetDevice: %s",
                                                  hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices)
        return lus::report_error(LUS_EINVAL, "lus_counting_sort_noalloc: device %d beyond %d", dev, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return lus::report_error(LUS_EHIP, "lus_counting_sort_noalloc: hipGetSymbolAddress: %s",
                                     hipGetErrorString(e));
        g_scratch_ptr[dev] = static_cast<int*>(p);
    }
    *out_dev = dev;
    *out = g_scratch_ptr[dev];
    return LUS_OK;
}

__global__ __launch_bounds__(lus::kBlock) void k_fill_zero(int* __restrict__ a, int m) {
    const int i = blockIdx.x * lus::kBlock + threadIdx.x;
    if (i < m) a[i] = 0;
}

}  // namespace

static_assert((size_t)LUS_NOALLOC_MAX_CELLS / 2048 + 2 <= kTilesInts, "scan tile bound");

extern "C" int lus_counting_sort_noalloc(const int* keys, int n, int num_cells, int* counts, int* sorted,
                                         void* stream) {
    if (n < 0 || num_cells <= 0 || !counts || (n > 0 && (!keys || !sorted)))
        return lus::report_error(LUS_EINVAL, "lus_counting_sort_noalloc: bad arguments");
    if (n > LUS_NOALLOC_MAX_KEYS || num_cells > LUS_NOALLOC_MAX_CELLS)
        return lus::report_error(LUS_EINVAL,
                                 "lus_counting_sort_noalloc: %d keys / %d cells exceed the static scratch "
                                 "(at most %d keys and %d cells)",
                                 n, num_cells, LUS_NOALLOC_MAX_KEYS, LUS_NOALLOC_MAX_CELLS);
    if ((size_t)lus::scan_tile_count(num_cells + 1) + 1 > kTilesInts)
        return lus::report_error(LUS_EINVAL, "lus_counting_sort_noalloc: scan tiles exceed the scratch");
    int dev = -1;
    int* base = nullptr;
    int rc = scratch_for_current_device(&dev, &base);
    if (rc != LUS_OK) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    int* hist = base + kCountsOff;
    int* slot = base + kSlotOff;
    int* tmp = base + kTmpOff;
    int* ks = base + kKsOff;
    int* tiles = base + kTilesOff;
    const int C1 = num_cells + 1;
    lus::launch_kernel(k_fill_zero, dim3((C1 + lus::kBlock - 1) / lus::kBlock), dim3(lus::kBlock), 0, s, hist, C1);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && n) e = lus::launch_count_keys(keys, n, hist, slot, s);
    // counts <- exclusive scan of the histogram (first slot of each key,
    // counts[num_cells] = n); the scan clears the histogram behind it
    if (e == hipSuccess) e = lus::launch_exclusive_scan(hist, counts, C1, tiles, 1, s);
    if (e == hipSuccess && n) e = lus::launch_scatter(keys, slot, counts, n, tmp, s);
    if (e == hipSuccess && n) e = lus::launch_stable_rank(tmp, keys, counts, n, sorted, ks, s);
    if (e != hipSuccess)
        return lus::report_error(LUS_EHIP, "lus_counting_sort_noalloc: launch failed on device %d: %s", dev,
                                 hipGetErrorString(e));
    return LUS_OK;
}
