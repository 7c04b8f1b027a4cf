// lus_kernels.hip -- gfx950 kernels of the particle substep.
//
// Build flags (lustrine_amd/build.py): -ffp-contract=off and correctly rounded
// fp32 divide/sqrt, so each expression below rounds exactly like the CPU
// oracle's; packed-fp32 instruction selection is disabled.
//
// Layout in HBM: one float4 per particle per array (x, y, z, w), arrays in
// cell order after the per-step counting sort.  Cell c's particles are slots
// [cell_start[c], cell_start[c+1]).  With the reference's y-major cell index
// (src/neighbors/Utils.hpp:24-33) the three z-neighbour cells of a (y, x) row
// are contiguous, so the 27-cell stencil is 9 contiguous slot ranges.
//
// Launch shape: 256 threads (4 waves of 64) per block, one particle (or cell
// entry) per thread.  Grids are padded to a multiple of 8 blocks and block b
// is remapped to tile (b % 8) * (grid / 8) + b / 8: the hardware hands
// consecutive blocks to the 8 XCDs in turn, so this gives each XCD one
// contiguous eighth of the cell-ordered particles and keeps a particle's
// neighbours in the L2 of the XCD that gathers them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>

#include "lus_device.hpp"
#include "lus_internal.hpp"
#include "lus_launch.hpp"

namespace lus {

KernelTimes& kernel_times() {
    static thread_local KernelTimes t;
    return t;
}

// ------------------------------------------------------------------ scan
// Two passes over tiles of kV * 256 ints (the sort tile: 1024, 2048 or 4096
// cells; kScanTileCells by default): k_scan_reduce writes each tile's sum,
// k_scan_down rescans the tile and adds the sum of all earlier tiles.  Up to
// kScanFuseTiles tiles every down block sums the earlier tile sums itself (at
// most 16 KB, L2 resident); above that k_scan_tiles scans them first in one
// block.  Each thread owns kV consecutive ints, moved as 16-byte accesses when
// the arrays are 16-byte aligned.
constexpr int kScanFuseTiles = 4096;

int scan_tile_count(int m, int tile) { return (m + tile - 1) / tile; }

bool scan_tile_valid(int tile) { return tile == 1024 || tile == 2048 || tile == 4096; }

// exclusive scan of one value per thread across the block (any multiple of 64
// threads up to 1024); *total receives the block sum
__device__ __forceinline__ int block_excl_scan(int v, int* lds, int* total) {
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int nw = blockDim.x >> 6;
    int incl = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        int t = __shfl_up(incl, d, 64);
        if (lane >= d) incl += t;
    }
    if (lane == 63) lds[w] = incl;
    __syncthreads();
    int woff = 0, tot = 0;
    for (int k = 0; k < nw; ++k) {
        int s = lds[k];
        woff += (k < w) ? s : 0;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return woff + incl - v;
}

template <bool kVec, int kV>
__device__ __forceinline__ void scan_load(const int* __restrict__ in, int base, int m, int v[kV]) {
    if (kVec && base + kV <= m) {
#pragma unroll
        for (int q = 0; q < kV; q += 4) {
            const int4 a = *reinterpret_cast<const int4*>(in + base + q);
            v[q] = a.x; v[q + 1] = a.y; v[q + 2] = a.z; v[q + 3] = a.w;
        }
    } else {
#pragma unroll
        for (int k = 0; k < kV; ++k) v[k] = (base + k < m) ? in[base + k] : 0;
    }
}

template <bool kVec, int kV>
__global__ __launch_bounds__(kBlock) void k_scan_reduce(const int* __restrict__ in, int m,
                                                         int* __restrict__ tile_sums) {
    __shared__ int lds[16];
    const int base = blockIdx.x * (kBlock * kV) + threadIdx.x * kV;
    int v[kV];
    scan_load<kVec, kV>(in, base, m, v);
    int s = 0;
#pragma unroll
    for (int k = 0; k < kV; ++k) s += v[k];
    int tot;
    block_excl_scan(s, lds, &tot);
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = tot;
}

// one block: exclusive scan of the tile sums in place; tile_sums[nt] = total
__global__ __launch_bounds__(1024) void k_scan_tiles(int* __restrict__ tile_sums, int nt) {
    __shared__ int lds[16];
    int carry = 0;
    for (int base = 0; base < nt; base += blockDim.x) {
        int k = base + threadIdx.x;
        int v = (k < nt) ? tile_sums[k] : 0;
        int tot;
        int e = block_excl_scan(v, lds, &tot);
        if (k < nt) tile_sums[k] = carry + e;
        carry += tot;
    }
    if (threadIdx.x == 0) tile_sums[nt] = carry;
}

// kFuse: tile_sums holds plain tile sums and the block adds those of tiles
// 0..blockIdx.x-1; otherwise tile_sums holds their exclusive scan.
template <bool kVec, bool kFuse, int kV>
__global__ __launch_bounds__(kBlock) void k_scan_down(int* in, int m, const int* __restrict__ tile_sums,
                                                       int* out, int zero_in) {
    __shared__ int lds[16];
    __shared__ int pre_lds[kBlock / 64];
    const int base = blockIdx.x * (kBlock * kV) + threadIdx.x * kV;
    int v[kV];
    scan_load<kVec, kV>(in, base, m, v);
    int s = 0;
#pragma unroll
    for (int k = 0; k < kV; ++k) s += v[k];
    int pre;
    if (kFuse) {
        int t = 0;
        for (int q = threadIdx.x; q < (int)blockIdx.x; q += kBlock) t += tile_sums[q];
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) t += __shfl_xor(t, d, 64);
        if ((threadIdx.x & 63) == 0) pre_lds[threadIdx.x >> 6] = t;
        // block_excl_scan's first barrier publishes pre_lds
    }
    int tot;
    int e = block_excl_scan(s, lds, &tot);
    if (kFuse) {
        pre = 0;
#pragma unroll
        for (int w = 0; w < kBlock / 64; ++w) pre += pre_lds[w];
    } else {
        pre = tile_sums[blockIdx.x];
    }
    e += pre;
    int o[kV];
#pragma unroll
    for (int k = 0; k < kV; ++k) {
        o[k] = e;
        e += v[k];
    }
    if (kVec && base + kV <= m) {
#pragma unroll
        for (int q = 0; q < kV; q += 4) {
            *reinterpret_cast<int4*>(out + base + q) = make_int4(o[q], o[q + 1], o[q + 2], o[q + 3]);
            if (zero_in) *reinterpret_cast<int4*>(in + base + q) = make_int4(0, 0, 0, 0);
        }
    } else {
#pragma unroll
        for (int k = 0; k < kV; ++k) {
            if (base + k < m) {
                out[base + k] = o[k];
                if (zero_in) in[base + k] = 0;
            }
        }
    }
}

template <bool kVec, int kV>
static void scan_launch(int* in, int* out, int m, int* tile_sums, int zero_in, hipStream_t s) {
    const int nt = scan_tile_count(m, kBlock * kV);
    launch_kernel((k_scan_reduce<kVec, kV>), dim3(nt), dim3(kBlock), 0, s, in, m, tile_sums);
    if (nt <= kScanFuseTiles) {
        launch_kernel((k_scan_down<kVec, true, kV>), dim3(nt), dim3(kBlock), 0, s, in, m,
                           tile_sums, out, zero_in);
    } else {
        launch_kernel(k_scan_tiles, dim3(1), dim3(1024), 0, s, tile_sums, nt);
        launch_kernel((k_scan_down<kVec, false, kV>), dim3(nt), dim3(kBlock), 0, s, in, m,
                           tile_sums, out, zero_in);
    }
}

template <bool kVec>
static void scan_launch_tile(int* in, int* out, int m, int* tile_sums, int zero_in, hipStream_t s, int tile) {
    if (tile == 1024) scan_launch<kVec, 4>(in, out, m, tile_sums, zero_in, s);
    else if (tile == 4096) scan_launch<kVec, 16>(in, out, m, tile_sums, zero_in, s);
    else scan_launch<kVec, 8>(in, out, m, tile_sums, zero_in, s);
}

hipError_t launch_exclusive_scan(const int* in, int* out, int m, int* tile_sums, int zero_in,
                                 hipStream_t s, int tile) {
    if (m <= 0) return hipSuccess;
    if (!scan_tile_valid(tile)) return hipErrorInvalidValue;
    // out may alias in (only when zero_in is 0: every thread reads its ints
    // before it writes them)
    const bool vec = ((reinterpret_cast<uintptr_t>(in) | reinterpret_cast<uintptr_t>(out)) & 15) == 0;
    if (vec) scan_launch_tile<true>(const_cast<int*>(in), out, m, tile_sums, zero_in, s, tile);
    else scan_launch_tile<false>(const_cast<int*>(in), out, m, tile_sums, zero_in, s, tile);
    return hipGetLastError();
}

// One-pass scan with decoupled look-back (the PBF step's cell histogram).
// Block b scans tile b, publishes the tile's sum in status[b] (bit 31 set:
// published), and the lanes of its first wave add the published sums of the
// earlier tiles, one tile per lane.  A sum not published after kLookbackSpins
// polls is computed by that lane from the tile's ints (`in` is not written by
// this kernel: the scatter that follows clears it), so no block waits on a
// block that has not started.  status holds this step's row, zeroed;
// status_next (nt_max ints) is the next step's row, zeroed here.
constexpr int kLookbackSpins = 1 << 16;
constexpr int kPublished = 1 << 31;

template <bool kVec, int kV>
__global__ __launch_bounds__(kBlock) void k_scan_lookback(const int* __restrict__ in, int m, int* __restrict__ out,
                                                           int* __restrict__ status, int* __restrict__ status_next,
                                                           int nt_max) {
    __shared__ int lds[16];
    __shared__ int prefix;
    constexpr int kTile = kBlock * kV;
    const int b = blockIdx.x;
    for (int q = b * kBlock + (int)threadIdx.x; q < nt_max; q += gridDim.x * kBlock) status_next[q] = 0;
    const int base = b * kTile + threadIdx.x * kV;
    int v[kV];
    scan_load<kVec, kV>(in, base, m, v);
    int s = 0;
#pragma unroll
    for (int k = 0; k < kV; ++k) s += v[k];
    int tot;
    int e = block_excl_scan(s, lds, &tot);
    if (threadIdx.x == 0) __hip_atomic_store(status + b, tot | kPublished, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (threadIdx.x < 64) {
        const int lane = (int)threadIdx.x;
        int acc = 0;
        for (int q = b - 1 - lane; q >= 0; q -= 64) {
            int st = __hip_atomic_load(status + q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            for (int spins = 0; (st & kPublished) == 0 && spins < kLookbackSpins; ++spins) {
                __builtin_amdgcn_s_sleep(1);
                st = __hip_atomic_load(status + q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            if (st & kPublished) {
                acc += st & ~kPublished;
            } else {
                const int hi = (q + 1) * kTile < m ? (q + 1) * kTile : m;
                for (int c = q * kTile; c < hi; ++c) acc += in[c];
            }
        }
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) acc += __shfl_xor(acc, d, 64);
        if (lane == 0) prefix = acc;
    }
    __syncthreads();
    e += prefix;
    int o[kV];
#pragma unroll
    for (int k = 0; k < kV; ++k) {
        o[k] = e;
        e += v[k];
    }
    if (kVec && base + kV <= m) {
#pragma unroll
        for (int q = 0; q < kV; q += 4)
            *reinterpret_cast<int4*>(out + base + q) = make_int4(o[q], o[q + 1], o[q + 2], o[q + 3]);
    } else {
#pragma unroll
        for (int k = 0; k < kV; ++k)
            if (base + k < m) out[base + k] = o[k];
    }
}

int scan_lookback_rows(int m) { return scan_tile_count(m, kScanTileMin); }

hipError_t launch_scan_lookback(const int* in, int* out, int m, int* status, int* status_next, int tile,
                                hipStream_t s) {
    if (m <= 0) return hipSuccess;
    if (!scan_tile_valid(tile) || in == out) return hipErrorInvalidValue;
    const int nt = scan_tile_count(m, tile), nt_max = scan_lookback_rows(m);
    const bool vec = ((reinterpret_cast<uintptr_t>(in) | reinterpret_cast<uintptr_t>(out)) & 15) == 0;
#define LUS_LB(V, KV) launch_kernel((k_scan_lookback<V, KV>), dim3(nt), dim3(kBlock), 0, s, in, m, out, status, \
                                    status_next, nt_max)
    if (vec) {
        if (tile == 1024) LUS_LB(true, 4);
        else if (tile == 4096) LUS_LB(true, 16);
        else LUS_LB(true, 8);
    } else {
        if (tile == 1024) LUS_LB(false, 4);
        else if (tile == 4096) LUS_LB(false, 16);
        else LUS_LB(false, 8);
    }
#undef LUS_LB
    return hipGetLastError();
}

__global__ __launch_bounds__(kBlock) void k_zero_ints(int* __restrict__ a, int m) {
    for (int c = blockIdx.x * kBlock + (int)threadIdx.x; c < m; c += gridDim.x * kBlock) a[c] = 0;
}

hipError_t launch_zero_ints(int* a, int m, hipStream_t s) {
    if (m <= 0) return hipSuccess;
    const int g = std::min((m + kBlock - 1) / kBlock, 2048);
    launch_kernel(k_zero_ints, dim3(g), dim3(kBlock), 0, s, a, m);
    return hipGetLastError();
}

// ------------------------------------------------------------------ sort
__global__ __launch_bounds__(kBlock) void k_cell_keys(const float4* __restrict__ pos, int n, LusGrid g,
                                                       int* __restrict__ key) {
    const int i = xcd_index();
    if (i >= n) return;
    float4 p = pos[i];
    key[i] = cell_of(p.x, p.y, p.z, g.cell_size, g.gridX, g.gridZ);
}

hipError_t launch_cell_keys(const float4* pos, int n, LusGrid g, int* key, hipStream_t s) {
    launch_kernel(k_cell_keys, dim3(grid_for(n)), dim3(kBlock), 0, s, pos, n, g, key);
    return hipGetLastError();
}

__global__ __launch_bounds__(kBlock) void k_count_keys(const int* __restrict__ key, int n,
                                                        int* __restrict__ counts,
                                                        int* __restrict__ slot) {
    const int i = xcd_index();
    const bool active = i < n;
    const int c = active ? key[i] : 0;
    const int sl = wave_atomic_slot(counts, c, active);
    if (active) slot[i] = sl;
}

hipError_t launch_count_keys(const int* key, int n, int* counts, int* slot, hipStream_t s) {
    launch_kernel(k_count_keys, dim3(grid_for(n)), dim3(kBlock), 0, s, key, n, counts, slot);
    return hipGetLastError();
}

// clear (nullable): the histogram to zero, clear_n ints, for the next sort
__global__ __launch_bounds__(kBlock) void k_scatter(const int* __restrict__ key,
                                                     const int* __restrict__ slot,
                                                     const int* __restrict__ cell_start, int n,
                                                     int* __restrict__ tmp, int* __restrict__ clear,
                                                     int clear_n) {
    const int i = xcd_index();
    if (clear)
        for (int c = blockIdx.x * kBlock + (int)threadIdx.x; c < clear_n; c += gridDim.x * kBlock) clear[c] = 0;
    if (i >= n) return;
    tmp[cell_start[key[i]] + slot[i]] = i;
}

hipError_t launch_scatter(const int* key, const int* slot, const int* cell_start, int n, int* tmp,
                          hipStream_t s, int* clear, int clear_n) {
    launch_kernel(k_scatter, dim3(grid_for(n)), dim3(kBlock), 0, s, key, slot, cell_start, n,
                       tmp, clear, clear_n);
    return hipGetLastError();
}

__global__ __launch_bounds__(kBlock) void k_stable_rank(const int* __restrict__ tmp,
                                                         const int* __restrict__ key,
                                                         const int* __restrict__ cell_start, int n,
                                                         int* __restrict__ perm,
                                                         int* __restrict__ key_sorted) {
    const int p = xcd_index();
    if (p >= n) return;
    const int i = tmp[p];
    const int c = key[i];
    const int o = stable_slot(tmp, cell_start, c, i);
    perm[o] = i;
    key_sorted[o] = c;
}

hipError_t launch_stable_rank(const int* tmp, const int* key, const int* cell_start, int n,
                              int* perm, int* key_sorted, hipStream_t s) {
    launch_kernel(k_stable_rank, dim3(grid_for(n)), dim3(kBlock), 0, s, tmp, key, cell_start, n,
                       perm, key_sorted);
    return hipGetLastError();
}

__global__ __launch_bounds__(kBlock) void k_gather4(const float4* __restrict__ src,
                                                     const int* __restrict__ perm, int n,
                                                     float4* __restrict__ dst) {
    const int i = xcd_index();
    if (i >= n) return;
    dst[i] = src[perm[i]];
}

// AoS float[3] rows from float4: row q = ids ? ids[r] : r receives src[r]
// (the host's vec3 layout, written on the device so a download is one copy)
__global__ __launch_bounds__(kBlock) void k_pack3(const float4* __restrict__ src, const int* __restrict__ ids,
                                                   int n, float* __restrict__ dst) {
    const int r = xcd_index();
    if (r >= n) return;
    const float4 a = src[r];
    const int q = ids ? ids[r] : r;
    dst[3 * (size_t)q] = a.x;
    dst[3 * (size_t)q + 1] = a.y;
    dst[3 * (size_t)q + 2] = a.z;
}

hipError_t launch_pack3(const float4* src, const int* ids, int n, float* dst, hipStream_t s) {
    launch_kernel(k_pack3, dim3(grid_for(n)), dim3(kBlock), 0, s, src, ids, n, dst);
    return hipGetLastError();
}

hipError_t launch_gather4(const float4* src, const int* perm, int n, float4* dst, hipStream_t s) {
    launch_kernel(k_gather4, dim3(grid_for(n)), dim3(kBlock), 0, s, src, perm, n, dst);
    return hipGetLastError();
}

__global__ __launch_bounds__(kBlock) void k_gather1(const int* __restrict__ src,
                                                     const int* __restrict__ perm, int n,
                                                     int* __restrict__ dst) {
    const int i = xcd_index();
    if (i >= n) return;
    dst[i] = src[perm[i]];
}

hipError_t launch_gather1(const int* src, const int* perm, int n, int* dst, hipStream_t s) {
    launch_kernel(k_gather1, dim3(grid_for(n)), dim3(kBlock), 0, s, src, perm, n, dst);
    return hipGetLastError();
}

// plain float4 copy: the HBM ceiling of a read + write stream.  Each thread
// moves 4 float4 per block-tile (4 loads in flight before the stores).
__global__ __launch_bounds__(kBlock) void k_stream_copy(const float4* __restrict__ src,
                                                         float4* __restrict__ dst, size_t n4) {
    const size_t tile = (size_t)kBlock * 4;
    for (size_t b = (size_t)blockIdx.x * tile; b < n4; b += (size_t)gridDim.x * tile) {
        const size_t i0 = b + threadIdx.x;
        float4 v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (i0 + u * kBlock < n4) v[u] = src[i0 + u * kBlock];
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (i0 + u * kBlock < n4) dst[i0 + u * kBlock] = v[u];
    }
}

hipError_t launch_stream_copy(const float4* src, float4* dst, size_t n4, hipStream_t s) {
    size_t blocks = (n4 + (size_t)kBlock * 4 - 1) / ((size_t)kBlock * 4);
    if (blocks > 256 * 64) blocks = 256 * 64;
    launch_kernel(k_stream_copy, dim3((unsigned)blocks), dim3(kBlock), 0, s, src, dst, n4);
    return hipGetLastError();
}

// ------------------------------------------------------------------ fluid lists
// find_neighbors_uniform_grid (src/neighbors/Neighbors.cpp:366-450): for the
// particle with original slot i, neighbour cells in (dy, dx, dz) order, each
// cell's sand entries (ascending slot, = the stable sort order) then its solid
// entries (ascending slot); every entry with d^2 <= h^2, self included.
__global__ __launch_bounds__(kBlock) void k_fluid_nbr(
    const float4* __restrict__ sand, const int* __restrict__ skey, const int* __restrict__ scs,
    const int* __restrict__ sorig, int n, int sand_start, const float4* __restrict__ solid,
    const int* __restrict__ ocs, const int* __restrict__ oref, LusGrid g, float h2,
    const int* __restrict__ inv, const long* __restrict__ off, int* __restrict__ cnt,
    int* __restrict__ idx, int fill) {
    const int i = xcd_index();
    if (i >= n) return;
    const int p = inv[i];
    const float4 me = sand[p];
    const int c = skey[p];
    const int gXZ = g.gridX * g.gridZ;
    const int yy = c / gXZ;
    const int rm = c - yy * gXZ;
    const int xx = rm / g.gridZ;
    const int zz = rm - xx * g.gridZ;
    long o = fill ? off[i] : 0;
    int count = 0;
    for (int y = -1; y <= 1; ++y) {
        for (int x = -1; x <= 1; ++x) {
            for (int z = -1; z <= 1; ++z) {
                if (xx + x < 0 || xx + x >= g.gridX || yy + y < 0 || yy + y >= g.gridY ||
                    zz + z < 0 || zz + z >= g.gridZ)
                    continue;
                const int nc = (yy + y) * gXZ + (xx + x) * g.gridZ + (zz + z);
                for (int j = scs[nc]; j < scs[nc + 1]; ++j) {
                    float4 q = sand[j];
                    float dx = me.x - q.x, dy = me.y - q.y, dz = me.z - q.z;
                    float a = dx * dx;
                    float b = dy * dy;
                    float e = dz * dz;
                    if ((a + b) + e <= h2) {
                        if (fill) idx[o + count] = sand_start + sorig[j];
                        ++count;
                    }
                }
                if (solid) {
                    for (int j = ocs[nc]; j < ocs[nc + 1]; ++j) {
                        float4 q = solid[j];
                        float dx = me.x - q.x, dy = me.y - q.y, dz = me.z - q.z;
                        float a = dx * dx;
                        float b = dy * dy;
                        float e = dz * dz;
                        if ((a + b) + e <= h2) {
                            if (fill) idx[o + count] = oref[j];
                            ++count;
                        }
                    }
                }
            }
        }
    }
    if (!fill) cnt[i] = count;
}

hipError_t launch_fluid_nbr(const float4* sand_sorted, const int* sand_key, const int* sand_cs,
                            const int* sand_orig, int n_sand, int sand_start,
                            const float4* solid_sorted, const int* solid_cs, const int* solid_ref,
                            LusGrid g, float h2, const int* inv_perm, const long* off, int* cnt,
                            int* idx, int fill, hipStream_t s) {
    launch_kernel(k_fluid_nbr, dim3(grid_for(n_sand)), dim3(kBlock), 0, s, sand_sorted,
                       sand_key, sand_cs, sand_orig, n_sand, sand_start, solid_sorted, solid_cs,
                       solid_ref, g, h2, inv_perm, off, cnt, idx, fill);
    return hipGetLastError();
}

// ------------------------------------------------------------------ sand
// integration loop of simulate_sand, src/Simulate.cpp:188-213
// Predicted position of sand slot i with gravity and the attract / blow
// forces (src/Simulate.cpp:185-212); writes v, attracted, x*, key; returns
// the cell.
// simulate_sand_v3's integration (src/Simulate.cpp:541-556): gravity, the
// unbounded attract / blow pushes, no box clamp; the cell index is reported
// (*err) when it falls outside the grid, where the reference indexes outside
// its cell vector, and used as the reference uses it otherwise
__device__ __forceinline__ int sand_predict_v3(int i, float4* __restrict__ x, float4* __restrict__ v,
                                               const SandK& k, float dt, float4* __restrict__ xs,
                                               int* __restrict__ key, int* __restrict__ err) {
    const float4 p = x[i];
    float4 vi = v[i];
    const float w = k.w;
    vi.x = vi.x + k.g[0] * dt;
    vi.y = vi.y + k.g[1] * dt;
    vi.z = vi.z + k.g[2] * dt;
    if (k.attract || k.blow) {
        const float d[3] = {k.player[0] - p.x, k.player[1] - p.y, k.player[2] - p.z};
        const float inv = 1.0f / sqrtf(dot3(d[0], d[1], d[2], d[0], d[1], d[2]));
        float add[3];
        if (k.attract) {
            for (int c = 0; c < 3; ++c) {
                float t = d[c] * inv;
                t = t * k.r;
                t = t * 1000.0f;
                t = t * dt;
                t = t * w;
                add[c] = t;
            }
            vi.x = vi.x + add[0];
            vi.y = vi.y + add[1];
            vi.z = vi.z + add[2];
        }
        if (k.blow) {
            for (int c = 0; c < 3; ++c) {
                float t = -(d[c] * inv);
                t = t * k.r;
                t = t * 200.0f;
                t = t * w;
                add[c] = t;
            }
            vi.x = vi.x + add[0];
            vi.y = vi.y + add[1];
            vi.z = vi.z + add[2];
        }
    }
    float4 q;
    q.x = p.x + vi.x * dt;
    q.y = p.y + vi.y * dt;
    q.z = p.z + vi.z * dt;
    q.w = 0.0f;
    v[i] = vi;
    xs[i] = q;
    const float px = q.x / k.cs, py = q.y / k.cs, pz = q.z / k.cs;
    bool bad = !(fabsf(px) < 1.0e9f && fabsf(py) < 1.0e9f && fabsf(pz) < 1.0e9f);
    int c = 0;
    if (!bad) {
        const long long cl = (long long)(int)py * k.gXZ + (long long)(int)px * k.gZ + (long long)(int)pz;
        bad = cl < 0 || cl >= (long long)k.gXZ * k.gY;
        c = (int)cl;
    }
    if (bad) {
        atomicOr(err, 1);
        c = 0;
    }
    key[i] = c;
    return c;
}

__device__ __forceinline__ int sand_predict(int i, float4* __restrict__ x, float4* __restrict__ v,
                                            int* __restrict__ att, const SandK& k, float dt,
                                            float4* __restrict__ xs, int* __restrict__ key) {
    const float4 p = x[i];
    float4 vi = v[i];
    const float w = k.w;
    const bool credits = k.variant == 1;
    int a = att[i];
    // simulate_sand_credits: a particle flagged 2 falls at 1 (:361-364)
    if (credits && (a & 2)) vi = make_float4(0.0f, -1.0f, 0.0f, vi.w);
    vi.x = vi.x + k.g[0] * dt;
    vi.y = vi.y + k.g[1] * dt;
    vi.z = vi.z + k.g[2] * dt;
    if (k.prev_attract && !k.attract) a = credits ? (a & ~1) : 0;
    if (k.attract) {
        const float d0x = k.player[0] - p.x, d0y = k.player[1] - p.y, d0z = k.player[2] - p.z;
        if (sqrtf(dot3(d0x, d0y, d0z, d0x, d0y, d0z)) < k.attract_radius) a = credits ? ((a | 1) & ~2) : 1;
        if (credits ? (a & 1) : a) {
            const float ox = k.player[0] + 0.0f, oy = k.player[1] + 1.5f, oz = k.player[2] + 0.0f;
            const float d[3] = {ox - p.x, oy - p.y, oz - p.z};
            const float inv = 1.0f / sqrtf(dot3(d[0], d[1], d[2], d[0], d[1], d[2]));
            float add[3];
            for (int c = 0; c < 3; ++c) {
                float t = d[c] * inv;
                t = t * 1.0f;
                t = t * k.attract_coeff;
                t = t * k.r;
                t = t * dt;
                t = t * w;
                add[c] = t;
            }
            vi.x = vi.x + add[0];
            vi.y = vi.y + add[1];
            vi.z = vi.z + add[2];
        }
    }
    if (k.blow) {
        const float d[3] = {k.player[0] - p.x, k.player[1] - p.y, k.player[2] - p.z};
        const float l = sqrtf(dot3(d[0], d[1], d[2], d[0], d[1], d[2]));
        if (l < k.blow_radius) {
            const float inv = 1.0f / sqrtf(dot3(d[0], d[1], d[2], d[0], d[1], d[2]));
            const float kk = (l > k.blow_radius) ? 0.0f : (1.0f - l / k.blow_radius);
            float add[3];
            for (int c = 0; c < 3; ++c) {
                float t = -(d[c] * inv);
                t = t * kk;
                t = t * k.blow_coeff;
                t = t * k.r;
                t = t * w;
                add[c] = t;
            }
            vi.x = vi.x + add[0];
            vi.y = vi.y + add[1];
            vi.z = vi.z + add[2];
            if (credits) a &= ~2;
        }
    }
    float4 q;
    q.x = rclamp(p.x + vi.x * dt, k.r, k.hi[0]);
    q.y = rclamp(p.y + vi.y * dt, k.r, k.hi[1]);
    q.z = rclamp(p.z + vi.z * dt, k.r, k.hi[2]);
    q.w = 0.0f;
    v[i] = vi;
    att[i] = a;
    xs[i] = q;
    const int c = cell_of(q.x, q.y, q.z, k.cs, k.gX, k.gZ);
    key[i] = c;
    return c;
}

__global__ __launch_bounds__(kBlock) void k_sand_predict_key(float4* __restrict__ x,
                                                              float4* __restrict__ v,
                                                              int* __restrict__ att, int n, SandK k,
                                                              float dt, int* __restrict__ err,
                                                              float4* __restrict__ xs,
                                                              int* __restrict__ key,
                                                              int* __restrict__ slot,
                                                              int* __restrict__ counts) {
    const int i = xcd_index();
    const bool active = i < n;
    int c = 0;
    if (active)
        c = k.variant == 2 ? sand_predict_v3(i, x, v, k, dt, xs, key, err) : sand_predict(i, x, v, att, k, dt, xs, key);
    // one atomic per run of equal keys in the wave (every lane calls it)
    const int sl = wave_atomic_slot(counts, c, active);
    if (active) slot[i] = sl;
}

hipError_t launch_sand_predict_key(float4* x, float4* v, int* att, int n, SandK k, float dt, int* err,
                                   float4* xs, int* key, int* slot, int* counts, hipStream_t s) {
    launch_kernel(k_sand_predict_key, dim3(grid_for(n)), dim3(kBlock), 0, s, x, v, att, n, k,
                       dt, err, xs, key, slot, counts);
    return hipGetLastError();
}

// stable counting sort permutation of x, x*, v, attracted
// (src/neighbors/Neighbors.cpp:280-304)
__global__ __launch_bounds__(kBlock) void k_sand_permute(
    const int* __restrict__ tmp, const int* __restrict__ key, const int* __restrict__ cell_start,
    const float4* __restrict__ x, const float4* __restrict__ xs, const float4* __restrict__ v,
    const int* __restrict__ att, int n, float4* __restrict__ x2, float4* __restrict__ xs2,
    float4* __restrict__ v2, int* __restrict__ att2, int* __restrict__ perm,
    int* __restrict__ key2) {
    const int p = xcd_index();
    if (p >= n) return;
    const int i = tmp[p];
    const int c = key[i];
    const int o = stable_slot(tmp, cell_start, c, i);
    x2[o] = x[i];
    xs2[o] = xs[i];
    v2[o] = v[i];
    att2[o] = att[i];
    perm[o] = i;
    key2[o] = c;
}

hipError_t launch_sand_permute(const int* tmp, const int* key, const int* cell_start,
                               const float4* x, const float4* xs, const float4* v, const int* att,
                               int n, float4* x2, float4* xs2, float4* v2, int* att2, int* perm,
                               int* key2, hipStream_t s) {
    launch_kernel(k_sand_permute, dim3(grid_for(n)), dim3(kBlock), 0, s, tmp, key, cell_start,
                       x, xs, v, att, n, x2, xs2, v2, att2, perm, key2);
    return hipGetLastError();
}

// Neighbor lists of find_neighbors_uniform_grid_v1
// (src/neighbors/Neighbors.cpp:306-361), in the order the reference builds
// them: entries pushed by earlier slots k < i (the symmetric push of :351-352)
// come first, in ascending k, then slot i's own scan of the clamped 27 cells
// in (y, x, z) order, each cell holding its solids first (the cached static
// lists, :293) then its sand slots; j = i is pushed twice.
// Solid entries are stored as n + (solid sorted position).
// Wave-staged: the slots of a wave lie in a few (y, x) cell columns, and
// within a column their z windows ascend with the slot, so for each stencil
// row the union of the wave's windows is one slot range per column run, in
// the sand table and in the solid table.  The wave copies those ranges (x, y,
// z of each slot) into LDS with coalesced loads and its lanes test their
// candidates there, in the same order with the same arithmetic; a row whose
// ranges exceed kSandStage slots or kSandSegs runs is read from global memory.
constexpr int kSandStage = 256;
constexpr int kSandSegs = 4;

// The reference's neighbour test of the sand lists (squared length within h2)
__device__ __forceinline__ bool sand_within(const float4 me, float qx, float qy, float qz, const SandK& k) {
    const float dx = me.x - qx, dy = me.y - qy, dz = me.z - qz;
    return dot3(dx, dy, dz, dx, dy, dz) <= k.h2;
}

// Hit mask of the cnt (<= 32) staged candidates b[3 t .. 3 t + 2]: bit t set
// when candidate t is within h, from the sign of h2 - r2 (r2 as sand_within
// forms it; the rounded difference of two floats has the sign of their exact
// difference and is +0 when they are equal), shifted in from the last
// candidate to the first, so one funnel shift replaces compare and branch.
__device__ __forceinline__ unsigned int sand_hits32(const float4 me, const lds_float* __restrict__ b, int cnt,
                                                    float h2) {
    unsigned int o = ~0u;  // outside bits; ones above bit cnt - 1
    for (int t = cnt - 1; t >= 0; --t) {
        const lds_float* __restrict__ p = b + 3 * t;
        const float dx = me.x - p[0], dy = me.y - p[1], dz = me.z - p[2];
        o = __builtin_amdgcn_alignbit(o, __float_as_uint(h2 - dot3(dx, dy, dz, dx, dy, dz)), 31u);
    }
    return ~o;
}

// A slot's list entries as its scan finds them: part 1 (sand slots j < i,
// pushed by j's own scan in the reference) from o1, part 2 (the own scan)
// from o2; counted only when not filling
struct SandHits {
    int o1 = 0, o2 = 0, n1 = 0, n2 = 0;
    __device__ __forceinline__ void solid(int js, int n, int fill, int* __restrict__ idx) {
        if (fill) idx[o2 + n2] = n + js;
        ++n2;
    }
    __device__ __forceinline__ void sand(int j, int i, int fill, int* __restrict__ idx, const SandK& k) {
        if (j < i) {
            if (i < k.sym_limit) {
                if (fill) idx[o1 + n1] = j;
                ++n1;
            }
        } else {
            if (fill) idx[o2 + n2] = j;
            ++n2;
            if (j == i && i < k.sym_limit) {  // j = i is pushed twice
                if (fill) idx[o2 + n2] = j;
                ++n2;
            }
        }
    }
};

__global__ __launch_bounds__(kBlock) void k_sand_nbr(const float4* __restrict__ xs,
                                                      const int* __restrict__ key2,
                                                      const int* __restrict__ scs, int n,
                                                      const float4* __restrict__ solid,
                                                      const int* __restrict__ ocs, SandK k,
                                                      const int* __restrict__ off,
                                                      int* __restrict__ cnt,
                                                      int* __restrict__ cnt1,
                                                      int* __restrict__ idx, int fill,
                                                      long cap, int* __restrict__ fail) {
    // per wave: the staged sand and solid ranges of a row
    __shared__ float stage[kBlock / 64][2][3 * kSandStage];
    // a fixed index buffer (fail non-null): a total past its capacity (or
    // past the int range) writes nothing and marks the step failed (uniform)
    if (fill && fail) {
        const int tot = off[n];
        if (tot < 0 || (long)tot > cap) {
            if (xcd_index() == 0) *fail = 1;
            return;
        }
    }
    const int i = xcd_index();
    const int lane = (int)(threadIdx.x & 63);
    // every lane of the wave takes part in the staging; a lane past n has no
    // rows and writes nothing
    const bool act = i < n;
    const int ii = act ? i : n - 1;
    lds_float* __restrict__ SS = (lds_float*)stage[threadIdx.x >> 6][0];
    lds_float* __restrict__ SO = (lds_float*)stage[threadIdx.x >> 6][1];
    const float4 me = xs[ii];
    const int c = key2[ii];
    const int yy = c / k.gXZ;
    const int tm = c - yy * k.gXZ;
    const int zz = tm % k.gZ;
    const int xx = tm / k.gZ;
    int yl = -1, yu = 1, xl = -1, xu = 1, zl = -1, zu = 1;
    if (yy + yl < 0) yl = 0;
    if (yy + yu >= k.gY) yu = 0;
    if (xx + xl < 0) xl = 0;
    if (xx + xu >= k.gX) xu = 0;
    if (zz + zl < 0) zl = 0;
    if (zz + zu >= k.gZ) zu = 0;
    const int nz = zu - zl + 1;  // z cells of a row (1 to 3)
    SandHits h;
    h.o1 = (fill && act) ? off[i] : 0;            // part 1: sand k < i
    h.o2 = (fill && act) ? off[i] + cnt1[i] : 0;  // part 2: own scan
    const unsigned long long below = (1ull << lane) - 1ull;  // lanes before this one
#pragma unroll 1
    for (int r = 0; r < 9; ++r) {
        const int y = r / 3 - 1, x = r % 3 - 1;
        const bool valid = act && y >= yl && y <= yu && x >= xl && x <= xu;
        // the row's z cells are consecutive: their 2 to 4 boundaries in
        // both cell tables are requested together
        const int row = valid ? (yy + y) * k.gXZ + (xx + x) * k.gZ + zz : 0;
        int sb[4], ob[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const bool in = valid && q <= nz;
            sb[q] = in ? scs[row + zl + q] : 0;
            ob[q] = (in && solid) ? ocs[row + zl + q] : 0;
        }
        const int s_hi = nz == 1 ? sb[1] : (nz == 2 ? sb[2] : sb[3]);
        const int o_hi = nz == 1 ? ob[1] : (nz == 2 ? ob[2] : ob[3]);
        // runs of lanes whose row lies in one cell column (their windows
        // ascend with the slot)
        const int col = valid ? (yy + y) * k.gX + (xx + x) : -1;
        const int pcol = __shfl_up(col, 1, 64);
        const int ncol = __shfl_down(col, 1, 64);
        const unsigned long long fm = __ballot(valid && (lane == 0 || pcol != col));
        const unsigned long long lm = __ballot(valid && (lane == 63 || ncol != col));
        const int nseg = __popcll(fm);
        if (nseg == 0) continue;  // uniform: no lane has this row
        // the runs' sand and solid ranges, packed in LDS at offsets s_off / o_off
        int s_start[kSandSegs], s_off[kSandSegs + 1], o_start[kSandSegs], o_off[kSandSegs + 1];
        int tot_s = 0, tot_o = 0;
        {
            unsigned long long fq = fm, lq = lm;
#pragma unroll
            for (int g = 0; g < kSandSegs; ++g) {
                s_off[g] = tot_s;
                o_off[g] = tot_o;
                s_start[g] = 0;
                o_start[g] = 0;
                if (g < nseg) {
                    const int lf = __ffsll((long long)fq) - 1, ll = __ffsll((long long)lq) - 1;
                    fq &= fq - 1ull;
                    lq &= lq - 1ull;
                    s_start[g] = __builtin_amdgcn_readlane(sb[0], lf);
                    tot_s += __builtin_amdgcn_readlane(s_hi, ll) - s_start[g];
                    o_start[g] = __builtin_amdgcn_readlane(ob[0], lf);
                    tot_o += __builtin_amdgcn_readlane(o_hi, ll) - o_start[g];
                }
            }
            s_off[kSandSegs] = tot_s;
            o_off[kSandSegs] = tot_o;
        }
        // uniform
        const bool staged = nseg <= kSandSegs && tot_s <= kSandStage && tot_o <= kSandStage;
        if (staged) {
            for (int q = lane; q < tot_s; q += 64) {
                int j = q - s_off[0] + s_start[0];
#pragma unroll
                for (int g = 1; g < kSandSegs; ++g)
                    if (q >= s_off[g]) j = q - s_off[g] + s_start[g];
                const float4 p = xs[j];
                SS[3 * q] = p.x;
                SS[3 * q + 1] = p.y;
                SS[3 * q + 2] = p.z;
            }
            for (int q = lane; q < tot_o; q += 64) {
                int j = q - o_off[0] + o_start[0];
#pragma unroll
                for (int g = 1; g < kSandSegs; ++g)
                    if (q >= o_off[g]) j = q - o_off[g] + o_start[g];
                const float4 p = solid[j];
                SO[3 * q] = p.x;
                SO[3 * q + 1] = p.y;
                SO[3 * q + 2] = p.z;
            }
            wave_lds_sync();
            // this lane's run: LDS place of slot j is j + ls (sand), j + lo (solid)
            const int seg = __popcll(fm & (below | (1ull << lane))) - 1;
            int ls = s_off[0] - s_start[0], lo = o_off[0] - o_start[0];
#pragma unroll
            for (int g = 1; g < kSandSegs; ++g)
                if (seg == g) {
                    ls = s_off[g] - s_start[g];
                    lo = o_off[g] - o_start[g];
                }
            if (valid) {
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    if (q >= nz) break;
                    // each cell's solids then sand slots, 32 candidates a mask
                    for (int t0 = ob[q]; t0 < ob[q + 1]; t0 += 32) {
                        const int cn = ob[q + 1] - t0 < 32 ? ob[q + 1] - t0 : 32;
                        unsigned int m = sand_hits32(me, SO + 3 * (t0 + lo), cn, k.h2);
                        if (!fill) {
                            h.n2 += __popc(m);
                            continue;
                        }
                        for (; m; m &= m - 1u) h.solid(t0 + __builtin_ctz(m), n, fill, idx);
                    }
                    for (int t0 = sb[q]; t0 < sb[q + 1]; t0 += 32) {
                        const int cn = sb[q + 1] - t0 < 32 ? sb[q + 1] - t0 : 32;
                        unsigned int m = sand_hits32(me, SS + 3 * (t0 + ls), cn, k.h2);
                        if (!fill) {
                            // part 1: slots below i (if i receives pushes); part
                            // 2: the rest, and i itself once more
                            const int below_i = i - t0;  // candidates of this mask below slot i
                            const unsigned int lowm =
                                below_i <= 0 ? 0u : (below_i >= 32 ? ~0u : (1u << below_i) - 1u);
                            const bool sym = i < k.sym_limit;
                            h.n1 += sym ? __popc(m & lowm) : 0;
                            h.n2 += __popc(m & ~lowm);
                            if (sym && below_i >= 0 && below_i < 32 && ((m >> below_i) & 1u)) ++h.n2;
                            continue;
                        }
                        for (; m; m &= m - 1u) h.sand(t0 + __builtin_ctz(m), i, fill, idx, k);
                    }
                }
            }
            wave_lds_sync();  // the next row's staging overwrites this one's
        } else if (valid) {
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                if (q >= nz) break;
                // candidates four at a time: the four loads are requested
                // together, then tested in slot order (the last candidate of
                // the cell stands in for places past its end)
                if (solid) {
                    const int je = ob[q + 1];
                    for (int j0 = ob[q]; j0 < je; j0 += 4) {
                        float4 p4[4];
#pragma unroll
                        for (int u = 0; u < 4; ++u) p4[u] = solid[j0 + u < je ? j0 + u : je - 1];
#pragma unroll
                        for (int u = 0; u < 4; ++u)
                            if (j0 + u < je && sand_within(me, p4[u].x, p4[u].y, p4[u].z, k)) h.solid(j0 + u, n, fill, idx);
                    }
                }
                const int je = sb[q + 1];
                for (int j0 = sb[q]; j0 < je; j0 += 4) {
                    float4 p4[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) p4[u] = xs[j0 + u < je ? j0 + u : je - 1];
#pragma unroll
                    for (int u = 0; u < 4; ++u)
                        if (j0 + u < je && sand_within(me, p4[u].x, p4[u].y, p4[u].z, k)) h.sand(j0 + u, i, fill, idx, k);
                }
            }
        }
    }
    if (!fill && act) {
        cnt[i] = h.n1 + h.n2;
        cnt1[i] = h.n1;
    }
}

hipError_t launch_sand_nbr(const float4* xs, const int* key2, const int* sand_cs, int n,
                           const float4* solid, const int* solid_cs, SandK k, const int* off,
                           int* cnt, int* cnt1, int* idx, int fill, hipStream_t s, long cap, int* fail) {
    launch_kernel(k_sand_nbr, dim3(grid_for(n)), dim3(kBlock), 0, s, xs, key2, sand_cs, n, solid,
                       solid_cs, k, off, cnt, cnt1, idx, fill, cap, fail);
    return hipGetLastError();
}

// The contact + friction projection of particle (pi, xoi) against a neighbour
// at pj into dp (src/Simulate.cpp:240-300); xoj_of() gives the neighbour's
// old position, asked for only for a sand contact.
// Returns whether the pair is in contact (within the diameter).
template <class XoJ>
__device__ __forceinline__ bool sand_pair(const float4 pi, const float4 xoi, const float4 pj, XoJ&& xoj_of,
                                          bool sand, const SandK& k, float& dpx, float& dpy, float& dpz) {
    float ijx = pi.x - pj.x, ijy = pi.y - pj.y, ijz = pi.z - pj.z;
    // |ij| == 0 -> ij = (0, 1e-5, 0) (src/Simulate.cpp:236-239; not in
    // simulate_sand_v3); the length is computed again only then
    float len = sqrtf(dot3(ijx, ijy, ijz, ijx, ijy, ijz));
    if (len == 0.0f && k.variant != 2) {
        ijx = 0.0f;
        ijy = 0.00001f;
        ijz = 0.0f;
        len = sqrtf(dot3(ijx, ijy, ijz, ijx, ijy, ijz));
    }
    if (len > k.d) return false;
    float sc;
    if (sand) {
        sc = k.collision * k.mass_n;
        sc = sc / k.mass_sum;
        sc = sc * (len - k.d);
    } else {
        sc = k.collision * (len - k.d);
    }
    const float tx = (sc * ijx) / len;
    const float ty = (sc * ijy) / len;
    const float tz = (sc * ijz) / len;
    dpx = dpx - tx;
    dpy = dpy - ty;
    dpz = dpz - tz;
    const float d = sqrtf(dot3(tx, ty, tz, tx, ty, tz));
    float xidx = (pi.x - tx) - xoi.x;
    float xidy = (pi.y - ty) - xoi.y;
    float xidz = (pi.z - tz) - xoi.z;
    float xjdx, xjdy, xjdz, nx, ny, nz;
    if (sand) {
        const float4 xoj = xoj_of();
        xjdx = (pj.x + tx) - xoj.x;
        xjdy = (pj.y + ty) - xoj.y;
        xjdz = (pj.z + tz) - xoj.z;
        nx = (pi.x - tx) - (pj.x + tx);
        ny = (pi.y - ty) - (pj.y + ty);
        nz = (pi.z - tz) - (pj.z + tz);
    } else {
        xjdx = 0.0f;
        xjdy = 0.0f;
        xjdz = 0.0f;
        nx = (pi.x - tx) - pj.x;
        ny = (pi.y - ty) - pj.y;
        nz = (pi.z - tz) - pj.z;
    }
    const float inv = 1.0f / sqrtf(dot3(nx, ny, nz, nx, ny, nz));
    nx = nx * inv;
    ny = ny * inv;
    nz = nz * inv;
    const float rx = xidx - xjdx, ry = xidy - xjdy, rz = xidz - xjdz;
    const float rd = dot3(rx, ry, rz, nx, ny, nz);
    const float xtx = rx - rd * nx, xty = ry - rd * ny, xtz = rz - rd * nz;
    const float lt = sqrtf(dot3(xtx, xty, xtz, xtx, xty, xtz));
    if ((d * k.mu_s) > (lt + 1e-9f)) {
        dpx = dpx - k.friction * xtx;
        dpy = dpy - k.friction * xty;
        dpz = dpz - k.friction * xtz;
    } else {
        const float q = (k.mu_k * d) / (lt + 1e-9f);
        const float ratio = q < 1.0f ? q : 1.0f;
        dpx = dpx - (k.friction * xtx) * ratio;
        dpy = dpy - (k.friction * xty) * ratio;
        dpz = dpz - (k.friction * xtz) * ratio;
    }
    return true;
}

// simulate_sand_v3's six half-space projections at positions_tmp
// (src/Simulate.cpp:601-616 with solve_boundary_collision_constraint,
// :119-131), summed in its wall order
__device__ __forceinline__ void sand_walls_v3(const float4 q, const SandK& k, float& ox, float& oy, float& oz) {
    const float nrm[6][3] = {{1.0f, 0.0f, 0.0f}, {-1.0f, 0.0f, 0.0f}, {0.0f, 1.0f, 0.0f},
                             {0.0f, -1.0f, 0.0f}, {0.0f, 0.0f, 1.0f}, {0.0f, 0.0f, -1.0f}};
    const float org[6][3] = {{0.0f, 0.0f, 0.0f}, {k.dom[0], 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f},
                             {0.0f, k.dom[1], 0.0f}, {0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, k.dom[2]}};
    float dx = 0.0f, dy = 0.0f, dz = 0.0f;
#pragma unroll
    for (int w = 0; w < 6; ++w) {
        const float rx = q.x - org[w][0], ry = q.y - org[w][1], rz = q.z - org[w][2];
        const float C = dot3(nrm[w][0], nrm[w][1], nrm[w][2], rx, ry, rz) - k.r;
        float bx = 0.0f, by = 0.0f, bz = 0.0f;
        if (!(C >= 0.0f)) {
            const float sv = C / dot3(nrm[w][0], nrm[w][1], nrm[w][2], nrm[w][0], nrm[w][1], nrm[w][2]);
            bx = (-sv) * nrm[w][0];
            by = (-sv) * nrm[w][1];
            bz = (-sv) * nrm[w][2];
        }
        dx = dx + k.bcoef * bx;
        dy = dy + k.bcoef * by;
        dz = dz + k.bcoef * bz;
    }
    ox = dx;
    oy = dy;
    oz = dz;
}

// One Jacobi substep of contact + friction (src/Simulate.cpp:228-308).  The
// list is walked four entries at a time (their indices and positions all
// requested before the four are tested; places past the list end read slot
// i, in bounds and unused), collecting the possible contacts, which a second
// loop projects in list order.
//
// simulate_sand_credits: a particle flagged 2 loses the flag at its first
// contact (:463-469; the three rand() draws there are not used).
// simulate_sand_v3: its four substeps all start from the same positions_tmp
// and give the same bits, so one pass computes them; then the half-space
// projections of positions_tmp are added and nothing is clamped (:566-616).
constexpr int kSandContacts = 16;  // contacts of a lane held between its walk and its projection loop

__global__ __launch_bounds__(kBlock) void k_sand_substep(const float4* __restrict__ A,
                                                          const float4* __restrict__ xold,
                                                          const float4* __restrict__ solid,
                                                          const int* __restrict__ off,
                                                          const int* __restrict__ idx, int n,
                                                          SandK k, float4* __restrict__ B,
                                                          int* __restrict__ att,
                                                          const int* __restrict__ fail) {
    // the lane's contacts in list order: contact[u * kBlock + thread]
    __shared__ int contact[kSandContacts * kBlock];
    if (fail && *fail) return;  // the lists were not written (uniform)
    const int i = xcd_index();
    if (i >= n) return;
    const float4 pi = A[i];
    const float4 xoi = xold[i];
    float dpx = 0.0f, dpy = 0.0f, dpz = 0.0f;
    bool contact_any = false;
    const int t0 = off[i], t1 = off[i + 1];
    auto at = [&](int j) { return j < n ? A[j] : solid[j - n]; };
    // Most neighbours lie beyond d, but over a wave some lane has a contact
    // at almost every step of the walk: the walk only collects the
    // neighbours whose squared length can give a length within d (d2_far),
    // and a second loop, as long as the wave's largest contact count,
    // projects them in list order (the same sums in the same order).
    int nc = 0;
    for (int t = t0; t < t1; t += 4) {
        int j[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) j[u] = (t + u < t1) ? idx[t + u] : i;
        float4 pj[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) pj[u] = at(j[u]);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const float ijx = pi.x - pj[u].x, ijy = pi.y - pj[u].y, ijz = pi.z - pj[u].z;
            if (t + u < t1 && j[u] != i && dot3(ijx, ijy, ijz, ijx, ijy, ijz) <= k.d2_far) {
                if (nc < kSandContacts) contact[nc * kBlock + (int)threadIdx.x] = j[u];
                ++nc;
            }
        }
    }
    if (nc <= kSandContacts) {
        // four contacts at a time: slots, positions and old positions
        // requested together (places past nc read slot i, not projected)
        for (int u = 0; u < nc; u += 4) {
            int jj[4];
            float4 pj[4], xoj[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) jj[q] = u + q < nc ? contact[(u + q) * kBlock + (int)threadIdx.x] : i;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                pj[q] = at(jj[q]);
                xoj[q] = xold[jj[q] < n ? jj[q] : i];
            }
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if (u + q < nc)
                    contact_any |= sand_pair(pi, xoi, pj[q], [&] { return xoj[q]; }, jj[q] < n, k, dpx, dpy, dpz);
        }
    } else {
        // more than the buffer holds: every neighbour again
        for (int t = t0; t < t1; ++t) {
            const int j = idx[t];
            if (j != i)
                contact_any |= sand_pair(pi, xoi, at(j), [&] { return xold[j]; }, j < n, k, dpx, dpy, dpz);
        }
    }
    float4 r;
    if (k.variant == 2) {
        float wx, wy, wz;
        sand_walls_v3(pi, k, wx, wy, wz);
        r.x = (pi.x + dpx) + wx;
        r.y = (pi.y + dpy) + wy;
        r.z = (pi.z + dpz) + wz;
    } else {
        r.x = rclamp(pi.x + dpx, k.r, k.hi[0]);
        r.y = rclamp(pi.y + dpy, k.r, k.hi[1]);
        r.z = rclamp(pi.z + dpz, k.r, k.hi[2]);
    }
    r.w = 0.0f;
    B[i] = r;
    if (k.variant == 1 && contact_any) {
        const int a = att[i];
        if (a & 2) att[i] = a & ~2;
    }
}

hipError_t launch_sand_substep(const float4* A, const float4* xold, const float4* solid,
                               const int* off, const int* idx, int n, SandK k, float4* B, int* att,
                               hipStream_t s, const int* fail) {
    launch_kernel(k_sand_substep, dim3(grid_for(n)), dim3(kBlock), 0, s, A, xold, solid, off,
                       idx, n, k, B, att, fail);
    return hipGetLastError();
}

// v = (x* - x) / dt; x = x*  (src/Simulate.cpp:316-319)
__global__ __launch_bounds__(kBlock) void k_sand_vupdate(const float4* __restrict__ A,
                                                          float4* __restrict__ x,
                                                          float4* __restrict__ v, int n,
                                                          float dt) {
    const int i = xcd_index();
    if (i >= n) return;
    const float4 a = A[i];
    const float4 p = x[i];
    v[i] = make_float4((a.x - p.x) / dt, (a.y - p.y) / dt, (a.z - p.z) / dt, 0.0f);
    x[i] = a;
}

hipError_t launch_sand_vupdate(const float4* A, float4* xold_to_new, float4* vout, int n, float dt,
                               hipStream_t s) {
    launch_kernel(k_sand_vupdate, dim3(grid_for(n)), dim3(kBlock), 0, s, A, xold_to_new, vout,
                       n, dt);
    return hipGetLastError();
}

// ------------------------------------------------------------------ grid-state consumers
// Particles listed by the cells of a cell box (query_cell_num_particles,
// src/Lustrine.cpp:864-914).  A cell's list is its static solids, then its
// sand slots; with y-major cells the z run of one (y, x) row is contiguous in
// both tables, so one thread counts a whole row from four cell starts.
// all_ids: every listed id; otherwise only ids below sand_end (sand slots
// below it, solids whose reference slot is below it).
__global__ __launch_bounds__(kBlock) void k_cell_box_count(const int* __restrict__ cs,
                                                           const int* __restrict__ ocs,
                                                           const int* __restrict__ solid_ref, int4 lo,
                                                           int4 hi, int gX, int gZ, int sand_end,
                                                           int all_ids, int* __restrict__ out) {
    const int nx = hi.x - lo.x + 1;
    const int rows = (hi.y - lo.y + 1) * nx;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    int c = 0;
    if (t < rows) {
        const int y = lo.y + t / nx, x = lo.x + t % nx;
        const int row = (y * gX + x) * gZ;
        const int a = cs[row + lo.z], b = cs[row + hi.z + 1];
        const int sa = ocs[row + lo.z], sb = ocs[row + hi.z + 1];
        if (all_ids) {
            c = (b - a) + (sb - sa);
        } else {
            const int e = b < sand_end ? b : sand_end;
            c = e > a ? e - a : 0;
            for (int q = sa; q < sb; ++q) c += solid_ref[q] < sand_end ? 1 : 0;
        }
    }
    for (int d = 32; d >= 1; d >>= 1) c += __shfl_xor(c, d, 64);
    if ((threadIdx.x & 63) == 0 && c) atomicAdd(out, c);
}

hipError_t launch_cell_box_count(const int* cs, const int* ocs, const int* solid_ref, const int* lo,
                                 const int* hi, int gX, int gZ, int sand_end, int all_ids, int* out,
                                 hipStream_t s) {
    const int rows = (hi[1] - lo[1] + 1) * (hi[0] - lo[0] + 1);
    const int4 l = make_int4(lo[0], lo[1], lo[2], 0), h = make_int4(hi[0], hi[1], hi[2], 0);
    launch_kernel(k_cell_box_count, dim3((rows + kBlock - 1) / kBlock), dim3(kBlock), 0, s, cs, ocs,
                       solid_ref, l, h, gX, gZ, sand_end, all_ids, out);
    return hipGetLastError();
}

// out[2k], out[2k+1] = first and end slot of cell cells[k]
__global__ __launch_bounds__(kBlock) void k_cell_ranges(const int* __restrict__ cs,
                                                        const int* __restrict__ cells, int n,
                                                        int2* __restrict__ out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int c = cells[k];
    out[k] = make_int2(cs[c], cs[c + 1]);
}

hipError_t launch_cell_ranges(const int* cs, const int* cells, int n, int2* out, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    launch_kernel(k_cell_ranges, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, cs, cells, n,
                       out);
    return hipGetLastError();
}

// Slot moves a[dst[k]] = a_old[src[k]] for three float4 arrays and one int
// array: every source is read into the staging arrays before any slot is
// written (two launches), so moves may chain and overlap.
__global__ __launch_bounds__(kBlock) void k_moves_gather(const int2* __restrict__ mv, int m,
                                                         const float4* __restrict__ a0,
                                                         const float4* __restrict__ a1,
                                                         const float4* __restrict__ a2,
                                                         const int* __restrict__ a3,
                                                         float4* __restrict__ t0, float4* __restrict__ t1,
                                                         float4* __restrict__ t2, int* __restrict__ t3) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    const int src = mv[k].y;
    t0[k] = a0[src];
    t1[k] = a1[src];
    t2[k] = a2[src];
    t3[k] = a3[src];
}

__global__ __launch_bounds__(kBlock) void k_moves_scatter(const int2* __restrict__ mv, int m,
                                                          float4* __restrict__ a0, float4* __restrict__ a1,
                                                          float4* __restrict__ a2, int* __restrict__ a3,
                                                          const float4* __restrict__ t0,
                                                          const float4* __restrict__ t1,
                                                          const float4* __restrict__ t2,
                                                          const int* __restrict__ t3) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    const int dst = mv[k].x;
    a0[dst] = t0[k];
    a1[dst] = t1[k];
    a2[dst] = t2[k];
    a3[dst] = t3[k];
}

// Particles inside the player's collision box (Bullet coupling,
// src/BulletPhysics.cpp:602-606): |x - player| <= half on every axis.
__device__ __forceinline__ bool in_player_box(const float4 p, const float4 c, const float4 h) {
    return fabsf(p.x - c.x) <= h.x && fabsf(p.y - c.y) <= h.y && fabsf(p.z - c.z) <= h.z;
}

// Ordered compaction of the first k slots inside the box, in two passes over
// slots in plain order (block b holds slots [256 b, 256 b + 256)): the count
// per block, then, after an exclusive scan of the counts, each match's rank
// = its block's offset + the matches before it in the block.
__global__ __launch_bounds__(kBlock) void k_box_count(const float4* __restrict__ x, int n, float4 c, float4 h,
                                                      int* __restrict__ block_counts) {
    __shared__ int wc[kBlock / 64];
    const int i = blockIdx.x * kBlock + threadIdx.x;
    const bool in = i < n && in_player_box(x[i], c, h);
    const unsigned long long m = __ballot(in);
    if ((threadIdx.x & 63) == 0) wc[threadIdx.x >> 6] = __popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int w = 0; w < kBlock / 64; ++w) t += wc[w];
        block_counts[blockIdx.x] = t;
    }
}

__global__ __launch_bounds__(kBlock) void k_box_write(const float4* __restrict__ x, int n, float4 c, float4 h,
                                                      const int* __restrict__ block_off, int k,
                                                      int* __restrict__ slots, float4* __restrict__ pos) {
    __shared__ int wc[kBlock / 64];
    const int i = blockIdx.x * kBlock + threadIdx.x;
    const float4 p = i < n ? x[i] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const bool in = i < n && in_player_box(p, c, h);
    const unsigned long long m = __ballot(in);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) wc[w] = __popcll(m);
    __syncthreads();
    int before = block_off[blockIdx.x];
    for (int v = 0; v < w; ++v) before += wc[v];
    const int rank = before + __popcll(m & ((1ull << lane) - 1ull));
    if (in && rank < k) {
        slots[rank] = i;
        pos[rank] = p;
    }
}

hipError_t launch_box_count(const float4* x, int n, const float* c, const float* h, int* block_counts,
                            hipStream_t s) {
    const int nb = (n + kBlock - 1) / kBlock;
    if (nb == 0) return hipSuccess;
    launch_kernel(k_box_count, dim3(nb), dim3(kBlock), 0, s, x, n, make_float4(c[0], c[1], c[2], 0.0f),
                       make_float4(h[0], h[1], h[2], 0.0f), block_counts);
    return hipGetLastError();
}

hipError_t launch_box_write(const float4* x, int n, const float* c, const float* h, const int* block_off, int k,
                            int* slots, float4* pos, hipStream_t s) {
    const int nb = (n + kBlock - 1) / kBlock;
    if (nb == 0 || k <= 0) return hipSuccess;
    launch_kernel(k_box_write, dim3(nb), dim3(kBlock), 0, s, x, n, make_float4(c[0], c[1], c[2], 0.0f),
                       make_float4(h[0], h[1], h[2], 0.0f), block_off, k, slots, pos);
    return hipGetLastError();
}

// ------------------------------------------------------------------ scene construction
// init_chunk_from_grid (src/Lustrine.cpp:676-722).  Fine lattice point t, in
// the x-major order of the reference's loops over X*s x Y*s x Z*s points,
// belongs to grid cell (x/s, y/s, z/s) (cells indexed x*Y*Z + y*Z + z); it
// becomes a particle when the cell is occupied (and, with mask_out, does not
// hold exactly 1).  Ranks come from an exclusive scan of the flags.
__device__ __forceinline__ int chunk_cell(long t, int X, int Y, int Z, int sub, int* xf, int* yf, int* zf) {
    const long Zf = (long)Z * sub, Yf = (long)Y * sub;
    *zf = (int)(t % Zf);
    const long r = t / Zf;
    *yf = (int)(r % Yf);
    *xf = (int)(r / Yf);
    return (*xf / sub) * Y * Z + (*yf / sub) * Z + (*zf / sub);
}

__device__ __forceinline__ bool chunk_keep(int v, int mask_out) { return v != 0 && (!mask_out || v != 1); }

__global__ __launch_bounds__(kBlock) void k_chunk_flags(const int* __restrict__ cells, int X, int Y, int Z, int sub,
                                                        int mask_out, long F, int* __restrict__ flags) {
    const long t = (long)blockIdx.x * kBlock + threadIdx.x;
    if (t >= F) return;
    int xf, yf, zf;
    flags[t] = chunk_keep(cells[chunk_cell(t, X, Y, Z, sub, &xf, &yf, &zf)], mask_out) ? 1 : 0;
}

// position = ((x * cs) + grid position) + cs/2 per axis, each operation
// rounded on its own as the reference's vec3 arithmetic does
__global__ __launch_bounds__(kBlock) void k_chunk_write(const int* __restrict__ cells,
                                                        const float4* __restrict__ cell_colors, int X, int Y,
                                                        int Z, int sub, int mask_out, long F, float4 origin,
                                                        float cs, const int* __restrict__ rank,
                                                        float4* __restrict__ pos, float4* __restrict__ col) {
    const long t = (long)blockIdx.x * kBlock + threadIdx.x;
    if (t >= F) return;
    int xf, yf, zf;
    const int c = chunk_cell(t, X, Y, Z, sub, &xf, &yf, &zf);
    if (!chunk_keep(cells[c], mask_out)) return;
    const int o = rank[t];
    const float half = __fmul_rn(cs, 0.5f);
    pos[o] = make_float4(__fadd_rn(__fadd_rn(__fmul_rn((float)xf, cs), origin.x), half),
                         __fadd_rn(__fadd_rn(__fmul_rn((float)yf, cs), origin.y), half),
                         __fadd_rn(__fadd_rn(__fmul_rn((float)zf, cs), origin.z), half), 0.0f);
    if (col) col[o] = cell_colors[c];
}

hipError_t launch_chunk_flags(const int* cells, int X, int Y, int Z, int sub, int mask_out, long F, int* flags,
                              hipStream_t s) {
    if (F <= 0) return hipSuccess;
    launch_kernel(k_chunk_flags, dim3((unsigned)((F + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, cells, X, Y,
                       Z, sub, mask_out, F, flags);
    return hipGetLastError();
}

hipError_t launch_chunk_write(const int* cells, const float4* cell_colors, int X, int Y, int Z, int sub,
                              int mask_out, long F, const float* origin, float cs, const int* rank, float4* pos,
                              float4* col, hipStream_t s) {
    if (F <= 0) return hipSuccess;
    launch_kernel(k_chunk_write, dim3((unsigned)((F + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, cells,
                       cell_colors, X, Y, Z, sub, mask_out, F, make_float4(origin[0], origin[1], origin[2], 0.0f),
                       cs, rank, pos, col);
    return hipGetLastError();
}

// init_grid_from_magika_voxel (src/VoxelLoader.cpp:66-90) after parsing:
// grid cell (x, y, z) of an sx x sy x sz model takes voxel x + y sx + z sx sy
// (colour index, 0 empty) and the palette colour / 255; *occupied counts the
// set voxels (zeroed by the caller)
__global__ __launch_bounds__(kBlock) void k_grid_from_voxels(const unsigned char* __restrict__ vox, int sx, int sy,
                                                             int sz, const uchar4* __restrict__ palette,
                                                             int* __restrict__ cells, float4* __restrict__ colors,
                                                             int* __restrict__ occupied) {
    const long g = (long)blockIdx.x * kBlock + threadIdx.x;
    const long N = (long)sx * sy * sz;
    int hit = 0;
    if (g < N) {
        const int z = (int)(g % sz);
        const long r = g / sz;
        const int y = (int)(r % sy);
        const int x = (int)(r / sy);
        const int v = vox[(long)x + (long)y * sx + (long)z * sx * sy];
        cells[g] = v;
        float4 c = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (v) {
            const uchar4 p = palette[v];
            c = make_float4(__fdiv_rn((float)p.x, 255.0f), __fdiv_rn((float)p.y, 255.0f),
                            __fdiv_rn((float)p.z, 255.0f), __fdiv_rn((float)p.w, 255.0f));
            hit = 1;
        }
        colors[g] = c;
    }
    const unsigned long long m = __ballot(hit);
    if ((threadIdx.x & 63) == 0 && m) atomicAdd(occupied, __popcll(m));
}

hipError_t launch_grid_from_voxels(const unsigned char* vox, int sx, int sy, int sz, const uchar4* palette,
                                   int* cells, float4* colors, int* occupied, hipStream_t s) {
    const long N = (long)sx * sy * sz;
    if (N <= 0) return hipSuccess;
    launch_kernel(k_grid_from_voxels, dim3((unsigned)((N + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, vox, sx,
                       sy, sz, palette, cells, colors, occupied);
    return hipGetLastError();
}

hipError_t launch_slot_moves(const int2* mv, int m, float4* a0, float4* a1, float4* a2, int* a3,
                             float4* t0, float4* t1, float4* t2, int* t3, hipStream_t s) {
    if (m <= 0) return hipSuccess;
    const dim3 g((m + kBlock - 1) / kBlock), b(kBlock);
    launch_kernel(k_moves_gather, g, b, 0, s, mv, m, a0, a1, a2, a3, t0, t1, t2, t3);
    launch_kernel(k_moves_scatter, g, b, 0, s, mv, m, a0, a1, a2, a3, t0, t1, t2, t3);
    return hipGetLastError();
}

}  // namespace lus
