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

// Histogram slot of key c: one atomicAdd per distinct key of the wave instead
// of one per lane (particles arrive in the previous step's cell order, so a
// wave holds few distinct keys).  Every lane of the wave must call it.
__device__ __forceinline__ int wave_atomic_slot(int* __restrict__ counts, int c, bool active) {
    const int lane = threadIdx.x & 63;
    unsigned long long pending = __ballot(active);
    int slot = 0;
    while (pending) {
        const int leader = __ffsll((unsigned long long)pending) - 1;
        const int lk = __shfl(c, leader, 64);
        const bool mine = active && c == lk;
        const unsigned long long same = __ballot(mine);
        int base = 0;
        if (lane == leader) base = atomicAdd(&counts[lk], __popcll(same));
        base = __shfl(base, leader, 64);
        if (mine) slot = base + __popcll(same & ((1ull << lane) - 1ull));
        pending &= ~same;
    }
    return slot;
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

}  // namespace lus
