// lus_device.hpp -- device helpers shared by the gfx950 kernel files.
#pragma once

#include <hip/hip_runtime.h>

namespace lus {

constexpr int kBlock = 256;
constexpr int kXcd = 8;

// blocks for n threads at `block` threads per block, padded to a multiple of 8
static inline int grid_for(int n, int block = kBlock) {
    int b = (n + block - 1) / block;
    if (b < 1) b = 1;
    return ((b + kXcd - 1) / kXcd) * kXcd;
}

// Global index with an XCD-aware block remap: the hardware deals blocks
// round-robin to the 8 XCDs, so block b works on tile (b % 8) * (grid / 8) +
// b / 8 and each XCD walks one contiguous eighth of the array (grid padded to
// a multiple of 8).
__device__ __forceinline__ int xcd_index() {
    const int nb = gridDim.x;
    const int b = blockIdx.x;
    const int tile = (b % kXcd) * (nb / kXcd) + b / kXcd;
    return tile * (int)blockDim.x + (int)threadIdx.x;
}

__device__ __forceinline__ float rclamp(float x, float lo, float hi) {
    float t = (x < lo) ? lo : x;
    return (hi < t) ? hi : t;
}

// get_cell_id, src/neighbors/Utils.hpp:24-33
__device__ __forceinline__ int cell_of(float x, float y, float z, float cs, int gX, int gZ) {
    float px = x / cs;
    float py = y / cs;
    float pz = z / cs;
    return ((int)py) * gX * gZ + ((int)px) * gZ + ((int)pz);
}

// Histogram slot of key c: one atomicAdd per run of equal keys in consecutive
// active lanes instead of one per lane (particles arrive in the previous
// step's cell order, so a wave holds few runs).  The run heads issue their
// atomics in one instruction and the wave waits once; each lane then takes its
// head's base plus its rank inside the run.  A key split over several runs
// gets one disjoint range per run.  Every lane of the wave must call it.
__device__ __forceinline__ int wave_atomic_slot(int* __restrict__ counts, int c, bool active) {
    const int lane = threadIdx.x & 63;
    const unsigned long long act = __ballot(active);
    const int pc = __shfl_up(c, 1, 64);
    const bool pact = lane > 0 && ((act >> (lane - 1)) & 1ull);
    const bool head = active && (!pact || pc != c);
    const unsigned long long heads = __ballot(head);
    const unsigned long long below = (1ull << lane) - 1ull;  // lanes < lane
    const unsigned long long upto = (2ull << lane) - 1ull;   // lanes <= lane (all for 63)
    int base = 0;
    if (head) {
        const unsigned long long after = heads & ~upto;
        const unsigned long long end = after ? ((1ull << (__ffsll(after) - 1)) - 1ull) : ~0ull;
        base = atomicAdd(&counts[c], __popcll(act & ~below & end));
    }
    const unsigned long long mine = heads & upto;
    const int h = mine ? 63 - __clzll(mine) : lane;
    const int b = __shfl(base, h, 64);
    return b + __popcll(act & below & ~((1ull << h) - 1ull));
}

// float -> int preserving order, for atomicMax on floats
__device__ __forceinline__ int ordered(float f) {
    int b = __float_as_int(f);
    return b >= 0 ? b : b ^ 0x7FFFFFFF;
}

// The atomic slots order a cell's entries arbitrarily.  Entry p of cell c
// moves to cell_start[c] + #{q in cell c : tmp[q] < tmp[p]}: ascending
// original index inside the cell, i.e. the unique stable order that
// Sorting::counting_sort produces (src/neighbors/Sorting.cpp:11-33).
__device__ __forceinline__ int stable_slot(const int* __restrict__ tmp, const int* __restrict__ cell_start,
                                           int c, int i) {
    const int s = cell_start[c];
    const int e = cell_start[c + 1];
    int r = 0;
    for (int q = s; q < e; ++q) r += (tmp[q] < i) ? 1 : 0;
    return s + r;
}

__device__ __forceinline__ float dot3(float ax, float ay, float az, float bx, float by, float bz) {
    const float t0 = ax * bx;
    const float t1 = ay * by;
    const float t2 = az * bz;
    return (t0 + t1) + t2;
}

// LDS float, addressed with 32-bit offsets (a generic pointer into LDS would
// be indexed with 64-bit arithmetic)
typedef __attribute__((address_space(3))) float lds_float;

// LDS accesses of one wave complete, and are seen by all its lanes, before
// the wave goes on (staging -> tests, tests -> the next staging)
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

}  // namespace lus
