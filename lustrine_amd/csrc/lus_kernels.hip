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

#include <climits>

#include "lus_internal.hpp"

namespace lus {

constexpr int kBlock = 256;
constexpr int kXcd = 8;

static inline int grid_for(int n) {
    int b = (n + kBlock - 1) / kBlock;
    if (b < 1) b = 1;
    return ((b + kXcd - 1) / kXcd) * kXcd;
}

__device__ __forceinline__ int xcd_index() {
    const int nb = gridDim.x;
    const int b = blockIdx.x;
    const int tile = (b % kXcd) * (nb / kXcd) + b / kXcd;
    return tile * kBlock + threadIdx.x;
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

// float -> int preserving order, for atomicMax on floats
__device__ __forceinline__ int ordered(float f) {
    int b = __float_as_int(f);
    return b >= 0 ? b : b ^ 0x7FFFFFFF;
}

// ------------------------------------------------------------------ scan
constexpr int kScanV = 8;
constexpr int kScanTile = kBlock * kScanV;

int scan_tile_count(int m) { return (m + kScanTile - 1) / kScanTile; }

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

__global__ __launch_bounds__(kBlock) void k_scan_reduce(const int* __restrict__ in, int m,
                                                         int* __restrict__ tile_sums) {
    __shared__ int lds[16];
    const int base = blockIdx.x * kScanTile + threadIdx.x * kScanV;
    int s = 0;
#pragma unroll
    for (int k = 0; k < kScanV; ++k) s += (base + k < m) ? in[base + k] : 0;
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

__global__ __launch_bounds__(kBlock) void k_scan_down(int* in, int m, const int* __restrict__ tile_sums,
                                                       int* out, int zero_in) {
    __shared__ int lds[16];
    const int base = blockIdx.x * kScanTile + threadIdx.x * kScanV;
    int v[kScanV];
    int s = 0;
#pragma unroll
    for (int k = 0; k < kScanV; ++k) {
        v[k] = (base + k < m) ? in[base + k] : 0;
        s += v[k];
    }
    int tot;
    int e = block_excl_scan(s, lds, &tot) + tile_sums[blockIdx.x];
#pragma unroll
    for (int k = 0; k < kScanV; ++k) {
        if (base + k < m) {
            out[base + k] = e;
            if (zero_in) in[base + k] = 0;
        }
        e += v[k];
    }
}

hipError_t launch_exclusive_scan(const int* in, int* out, int m, int* tile_sums, int zero_in,
                                 hipStream_t s) {
    const int nt = scan_tile_count(m);
    if (nt > 0) {
        hipLaunchKernelGGL(k_scan_reduce, dim3(nt), dim3(kBlock), 0, s, in, m, tile_sums);
    }
    hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(1024), 0, s, tile_sums, nt);
    if (nt > 0) {
        hipLaunchKernelGGL(k_scan_down, dim3(nt), dim3(kBlock), 0, s, const_cast<int*>(in), m,
                           tile_sums, out, zero_in);
    }
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
    hipLaunchKernelGGL(k_cell_keys, dim3(grid_for(n)), dim3(kBlock), 0, s, pos, n, g, key);
    return hipGetLastError();
}

__global__ __launch_bounds__(kBlock) void k_count_keys(const int* __restrict__ key, int n,
                                                        int* __restrict__ counts,
                                                        int* __restrict__ slot) {
    const int i = xcd_index();
    if (i >= n) return;
    slot[i] = atomicAdd(&counts[key[i]], 1);
}

hipError_t launch_count_keys(const int* key, int n, int* counts, int* slot, hipStream_t s) {
    hipLaunchKernelGGL(k_count_keys, dim3(grid_for(n)), dim3(kBlock), 0, s, key, n, counts, slot);
    return hipGetLastError();
}

__global__ __launch_bounds__(kBlock) void k_scatter(const int* __restrict__ key,
                                                     const int* __restrict__ slot,
                                                     const int* __restrict__ cell_start, int n,
                                                     int* __restrict__ tmp) {
    const int i = xcd_index();
    if (i >= n) return;
    tmp[cell_start[key[i]] + slot[i]] = i;
}

hipError_t launch_scatter(const int* key, const int* slot, const int* cell_start, int n, int* tmp,
                          hipStream_t s) {
    hipLaunchKernelGGL(k_scatter, dim3(grid_for(n)), dim3(kBlock), 0, s, key, slot, cell_start, n,
                       tmp);
    return hipGetLastError();
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
    hipLaunchKernelGGL(k_stable_rank, dim3(grid_for(n)), dim3(kBlock), 0, s, tmp, key, cell_start, n,
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

hipError_t launch_gather4(const float4* src, const int* perm, int n, float4* dst, hipStream_t s) {
    hipLaunchKernelGGL(k_gather4, dim3(grid_for(n)), dim3(kBlock), 0, s, src, perm, n, dst);
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
    hipLaunchKernelGGL(k_gather1, dim3(grid_for(n)), dim3(kBlock), 0, s, src, perm, n, dst);
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
    hipLaunchKernelGGL(k_fluid_nbr, dim3(grid_for(n_sand)), dim3(kBlock), 0, s, sand_sorted,
                       sand_key, sand_cs, sand_orig, n_sand, sand_start, solid_sorted, solid_cs,
                       solid_ref, g, h2, inv_perm, off, cnt, idx, fill);
    return hipGetLastError();
}

// ------------------------------------------------------------------ PBF
__device__ __forceinline__ float4 pbf_predict(float4 x, float4 v, const PbfK& k, float dt) {
    float vx = v.x + k.g[0] * dt;
    float vy = v.y + k.g[1] * dt;
    float vz = v.z + k.g[2] * dt;
    float4 r;
    r.x = rclamp(x.x + vx * dt, k.lo[0], k.hi[0]);
    r.y = rclamp(x.y + vy * dt, k.lo[1], k.hi[1]);
    r.z = rclamp(x.z + vz * dt, k.lo[2], k.hi[2]);
    r.w = 0.0f;
    return r;
}

// 27-cell stencil as 9 contiguous slot ranges, ascending slot order
template <class F>
__device__ __forceinline__ void for_each_candidate(int key, const int* __restrict__ cs, const PbfK& k,
                                                   F&& f) {
    const int cy = key / k.gXZ;
    const int rm = key - cy * k.gXZ;
    const int cx = rm / k.gZ;
    const int cz = rm - cx * k.gZ;
    const int z0 = cz > 0 ? cz - 1 : 0;
    const int z1 = cz + 1 < k.gZ ? cz + 1 : k.gZ - 1;
    for (int dy = -1; dy <= 1; ++dy) {
        const int y = cy + dy;
        if (y < 0 || y >= k.gY) continue;
        for (int dx = -1; dx <= 1; ++dx) {
            const int x = cx + dx;
            if (x < 0 || x >= k.gX) continue;
            const int row = y * k.gXZ + x * k.gZ;
            const int j0 = cs[row + z0];
            const int j1 = cs[row + z1 + 1];
            for (int j = j0; j < j1; ++j) f(j);
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_pbf_key(const float4* __restrict__ x,
                                                     const float4* __restrict__ v, int n, PbfK k,
                                                     float dt, int* __restrict__ key,
                                                     int* __restrict__ slot,
                                                     int* __restrict__ counts,
                                                     int* __restrict__ maxc, int nmaxc) {
    const int i = xcd_index();
    if (i < nmaxc) maxc[i] = INT_MIN;
    if (i >= n) return;
    const float4 xs = pbf_predict(x[i], v[i], k, dt);
    const int c = cell_of(xs.x, xs.y, xs.z, k.cs, k.gX, k.gZ);
    key[i] = c;
    slot[i] = atomicAdd(&counts[c], 1);
}

hipError_t launch_pbf_key(const float4* x, const float4* v, int n, PbfK k, float dt, int* key,
                          int* slot, int* counts, int* maxc, int nmaxc, hipStream_t s) {
    hipLaunchKernelGGL(k_pbf_key, dim3(grid_for(n)), dim3(kBlock), 0, s, x, v, n, k, dt, key, slot,
                       counts, maxc, nmaxc);
    return hipGetLastError();
}

__global__ __launch_bounds__(kBlock) void k_pbf_permute(
    const int* __restrict__ tmp, const int* __restrict__ key, const int* __restrict__ cell_start,
    const float4* __restrict__ x, const float4* __restrict__ v, const int* __restrict__ id, int n,
    PbfK k, float dt, float4* __restrict__ x2, float4* __restrict__ xa, int* __restrict__ id2,
    int* __restrict__ key2) {
    const int p = xcd_index();
    if (p >= n) return;
    const int i = tmp[p];
    const int c = key[i];
    const int o = stable_slot(tmp, cell_start, c, i);
    const float4 xi = x[i];
    x2[o] = xi;
    xa[o] = pbf_predict(xi, v[i], k, dt);
    id2[o] = id[i];
    key2[o] = c;
}

hipError_t launch_pbf_permute(const int* tmp, const int* key, const int* cell_start,
                              const float4* x, const float4* v, const int* id, int n, PbfK k,
                              float dt, float4* x2, float4* xa, int* id2, int* key2,
                              hipStream_t s) {
    hipLaunchKernelGGL(k_pbf_permute, dim3(grid_for(n)), dim3(kBlock), 0, s, tmp, key, cell_start,
                       x, v, id, n, k, dt, x2, xa, id2, key2);
    return hipGetLastError();
}

// lambda pass: density constraint and its gradient norm (PBF eq. 1, 8, 11)
__global__ __launch_bounds__(kBlock) void k_pbf_lambda(const float4* __restrict__ xa,
                                                        const int* __restrict__ key2,
                                                        const int* __restrict__ cs, int n, PbfK k,
                                                        float4* __restrict__ xl,
                                                        int* __restrict__ maxc) {
    const int i = xcd_index();
    float Ci = -1e30f;
    if (i < n) {
        const float4 pi = xa[i];
        float rho = 0.0f, a = 0.0f, gx = 0.0f, gy = 0.0f, gz = 0.0f;
        for_each_candidate(key2[i], cs, k, [&](int j) {
            const float4 pj = xa[j];
            const float dx = pi.x - pj.x;
            const float dy = pi.y - pj.y;
            const float dz = pi.z - pj.z;
            const float r2 = dx * dx + dy * dy + dz * dz;
            if (r2 <= k.h2) {
                const float t = k.hf2 - k.f2 * r2;
                rho = rho + (t * t) * t;
                if (r2 > 0.0f) {
                    const float rl = sqrtf(r2);
                    const float g = k.hf - k.f * rl;
                    const float s = ((k.SPf * g) * g) / rl;
                    a = a + (s * s) * r2;
                    gx = gx + s * dx;
                    gy = gy + s * dy;
                    gz = gz + s * dz;
                }
            }
        });
        Ci = rho * k.c_rho - 1.0f;
        const float gg = gx * gx + gy * gy + gz * gz;
        const float den = k.mr0sq * (a + gg) + k.eps;
        xl[i] = make_float4(pi.x, pi.y, pi.z, -Ci / den);
    }
    // wave max, one atomic per wave
    int m = ordered(Ci);
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        int o = __shfl_xor(m, d, 64);
        m = o > m ? o : m;
    }
    if ((threadIdx.x & 63) == 0) atomicMax(maxc, m);
}

hipError_t launch_pbf_lambda(const float4* xa, const int* key2, const int* cs, int n, PbfK k,
                             float4* xl, int* maxc, hipStream_t s) {
    hipLaunchKernelGGL(k_pbf_lambda, dim3(grid_for(n)), dim3(kBlock), 0, s, xa, key2, cs, n, k, xl,
                       maxc);
    return hipGetLastError();
}

// delta-p pass with s_corr (PBF eq. 12-14), Jacobi update, box clamp
__global__ __launch_bounds__(kBlock) void k_pbf_deltap(const float4* __restrict__ xl,
                                                        const int* __restrict__ key2,
                                                        const int* __restrict__ cs, int n, PbfK k,
                                                        float4* __restrict__ xa,
                                                        const float4* __restrict__ xold,
                                                        float4* __restrict__ vout, float dt) {
    const int i = xcd_index();
    if (i >= n) return;
    const float4 pi = xl[i];
    const float li = pi.w;
    float ax = 0.0f, ay = 0.0f, az = 0.0f;
    for_each_candidate(key2[i], cs, k, [&](int j) {
        const float4 pj = xl[j];
        const float dx = pi.x - pj.x;
        const float dy = pi.y - pj.y;
        const float dz = pi.z - pj.z;
        const float r2 = dx * dx + dy * dy + dz * dz;
        if (r2 <= k.h2 && r2 > 0.0f) {
            const float rl = sqrtf(r2);
            const float g = k.hf - k.f * rl;
            const float s = ((k.SPf * g) * g) / rl;
            const float t = k.hf2 - k.f2 * r2;
            const float ratio = ((t * t) * t) * k.inv_tq3;
            float pw = ratio;
            for (int e = 1; e < k.n_scorr; ++e) pw = pw * ratio;
            const float sc = -k.scorr_k * pw;
            const float coef = (li + pj.w) + sc;
            const float c = coef * s;
            ax = ax + c * dx;
            ay = ay + c * dy;
            az = az + c * dz;
        }
    });
    float4 r;
    r.x = rclamp(pi.x + ax * k.neg_inv_r0, k.lo[0], k.hi[0]);
    r.y = rclamp(pi.y + ay * k.neg_inv_r0, k.lo[1], k.hi[1]);
    r.z = rclamp(pi.z + az * k.neg_inv_r0, k.lo[2], k.hi[2]);
    r.w = 0.0f;
    xa[i] = r;
    if (vout) {
        const float4 xo = xold[i];
        vout[i] = make_float4((r.x - xo.x) / dt, (r.y - xo.y) / dt, (r.z - xo.z) / dt, 0.0f);
    }
}

hipError_t launch_pbf_deltap(const float4* xl, const int* key2, const int* cs, int n, PbfK k,
                             float4* xa, const float4* xold, float4* vout, float dt,
                             hipStream_t s) {
    hipLaunchKernelGGL(k_pbf_deltap, dim3(grid_for(n)), dim3(kBlock), 0, s, xl, key2, cs, n, k, xa,
                       xold, vout, dt);
    return hipGetLastError();
}

__global__ __launch_bounds__(kBlock) void k_pbf_velocity(const float4* __restrict__ xa,
                                                          const float4* __restrict__ xold, int n,
                                                          float dt, float4* __restrict__ vout) {
    const int i = xcd_index();
    if (i >= n) return;
    const float4 r = xa[i];
    const float4 xo = xold[i];
    vout[i] = make_float4((r.x - xo.x) / dt, (r.y - xo.y) / dt, (r.z - xo.z) / dt, 0.0f);
}

hipError_t launch_pbf_velocity(const float4* xa, const float4* xold, int n, float dt,
                               float4* vout, hipStream_t s) {
    hipLaunchKernelGGL(k_pbf_velocity, dim3(grid_for(n)), dim3(kBlock), 0, s, xa, xold, n, dt, vout);
    return hipGetLastError();
}

// XSPH viscosity (PBF eq. 17) and commit x = x*
__global__ __launch_bounds__(kBlock) void k_pbf_xsph(const float4* __restrict__ xa,
                                                      const float4* __restrict__ vt,
                                                      const int* __restrict__ key2,
                                                      const int* __restrict__ cs, int n, PbfK k,
                                                      float4* __restrict__ vout,
                                                      float4* __restrict__ xout) {
    const int i = xcd_index();
    if (i >= n) return;
    const float4 pi = xa[i];
    const float4 vi = vt[i];
    float sx = 0.0f, sy = 0.0f, sz = 0.0f;
    for_each_candidate(key2[i], cs, k, [&](int j) {
        const float4 pj = xa[j];
        const float dx = pi.x - pj.x;
        const float dy = pi.y - pj.y;
        const float dz = pi.z - pj.z;
        const float r2 = dx * dx + dy * dy + dz * dz;
        if (r2 <= k.h2) {
            const float t = k.hf2 - k.f2 * r2;
            const float w = (t * t) * t;
            const float4 vj = vt[j];
            sx = sx + (vj.x - vi.x) * w;
            sy = sy + (vj.y - vi.y) * w;
            sz = sz + (vj.z - vi.z) * w;
        }
    });
    vout[i] = make_float4(vi.x + k.cW * sx, vi.y + k.cW * sy, vi.z + k.cW * sz, 0.0f);
    xout[i] = make_float4(pi.x, pi.y, pi.z, 0.0f);
}

hipError_t launch_pbf_xsph(const float4* xa, const float4* vt, const int* key2, const int* cs,
                           int n, PbfK k, float4* vout, float4* xout, hipStream_t s) {
    hipLaunchKernelGGL(k_pbf_xsph, dim3(grid_for(n)), dim3(kBlock), 0, s, xa, vt, key2, cs, n, k,
                       vout, xout);
    return hipGetLastError();
}

// ------------------------------------------------------------------ sand
__device__ __forceinline__ float dot3(float ax, float ay, float az, float bx, float by, float bz) {
    const float t0 = ax * bx;
    const float t1 = ay * by;
    const float t2 = az * bz;
    return (t0 + t1) + t2;
}

// integration loop of simulate_sand, src/Simulate.cpp:188-213
__global__ __launch_bounds__(kBlock) void k_sand_predict_key(float4* __restrict__ x,
                                                              float4* __restrict__ v,
                                                              int* __restrict__ att, int n, SandK k,
                                                              float dt, float4* __restrict__ xs,
                                                              int* __restrict__ key,
                                                              int* __restrict__ slot,
                                                              int* __restrict__ counts) {
    const int i = xcd_index();
    if (i >= n) return;
    const float4 p = x[i];
    float4 vi = v[i];
    const float w = k.w;
    vi.x = vi.x + k.g[0] * dt;
    vi.y = vi.y + k.g[1] * dt;
    vi.z = vi.z + k.g[2] * dt;
    int a = att[i];
    if (k.prev_attract && !k.attract) a = 0;
    if (k.attract) {
        const float d0x = k.player[0] - p.x, d0y = k.player[1] - p.y, d0z = k.player[2] - p.z;
        if (sqrtf(dot3(d0x, d0y, d0z, d0x, d0y, d0z)) < k.attract_radius) a = 1;
        if (a) {
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
    slot[i] = atomicAdd(&counts[c], 1);
}

hipError_t launch_sand_predict_key(float4* x, float4* v, int* att, int n, SandK k, float dt,
                                   float4* xs, int* key, int* slot, int* counts, hipStream_t s) {
    hipLaunchKernelGGL(k_sand_predict_key, dim3(grid_for(n)), dim3(kBlock), 0, s, x, v, att, n, k,
                       dt, xs, key, slot, counts);
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
    hipLaunchKernelGGL(k_sand_permute, dim3(grid_for(n)), dim3(kBlock), 0, s, tmp, key, cell_start,
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
__global__ __launch_bounds__(kBlock) void k_sand_nbr(const float4* __restrict__ xs,
                                                      const int* __restrict__ key2,
                                                      const int* __restrict__ scs, int n,
                                                      const float4* __restrict__ solid,
                                                      const int* __restrict__ ocs, SandK k,
                                                      const int* __restrict__ off,
                                                      int* __restrict__ cnt,
                                                      int* __restrict__ cnt1,
                                                      int* __restrict__ idx, int fill) {
    const int i = xcd_index();
    if (i >= n) return;
    const float4 me = xs[i];
    const int c = key2[i];
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
    int o1 = fill ? off[i] : 0;            // part 1: sand k < i
    int o2 = fill ? off[i] + cnt1[i] : 0;  // part 2: own scan
    int n1 = 0, n2 = 0;
    for (int y = yl; y <= yu; ++y) {
        for (int x = xl; x <= xu; ++x) {
            for (int z = zl; z <= zu; ++z) {
                const int nc = (yy + y) * k.gXZ + (xx + x) * k.gZ + (zz + z);
                if (solid) {
                    for (int j = ocs[nc]; j < ocs[nc + 1]; ++j) {
                        const float4 q = solid[j];
                        const float dx = me.x - q.x, dy = me.y - q.y, dz = me.z - q.z;
                        if (dot3(dx, dy, dz, dx, dy, dz) <= k.h2) {
                            if (fill) idx[o2 + n2] = n + j;
                            ++n2;
                        }
                    }
                }
                for (int j = scs[nc]; j < scs[nc + 1]; ++j) {
                    const float4 q = xs[j];
                    const float dx = me.x - q.x, dy = me.y - q.y, dz = me.z - q.z;
                    if (dot3(dx, dy, dz, dx, dy, dz) <= k.h2) {
                        if (j < i) {
                            // pushed into slot i by slot j's scan
                            if (i < k.sym_limit) {
                                if (fill) idx[o1 + n1] = j;
                                ++n1;
                            }
                        } else {
                            if (fill) idx[o2 + n2] = j;
                            ++n2;
                            if (j == i && i < k.sym_limit) {
                                if (fill) idx[o2 + n2] = j;
                                ++n2;
                            }
                        }
                    }
                }
            }
        }
    }
    if (!fill) {
        cnt[i] = n1 + n2;
        cnt1[i] = n1;
    }
}

hipError_t launch_sand_nbr(const float4* xs, const int* key2, const int* sand_cs, int n,
                           const float4* solid, const int* solid_cs, SandK k, const int* off,
                           int* cnt, int* cnt1, int* idx, int fill, hipStream_t s) {
    hipLaunchKernelGGL(k_sand_nbr, dim3(grid_for(n)), dim3(kBlock), 0, s, xs, key2, sand_cs, n, solid,
                       solid_cs, k, off, cnt, cnt1, idx, fill);
    return hipGetLastError();
}

// One Jacobi substep of contact + friction (src/Simulate.cpp:228-308)
__global__ __launch_bounds__(kBlock) void k_sand_substep(const float4* __restrict__ A,
                                                          const float4* __restrict__ xold,
                                                          const float4* __restrict__ solid,
                                                          const int* __restrict__ off,
                                                          const int* __restrict__ idx, int n,
                                                          SandK k, float4* __restrict__ B) {
    const int i = xcd_index();
    if (i >= n) return;
    const float4 pi = A[i];
    const float4 xoi = xold[i];
    float dpx = 0.0f, dpy = 0.0f, dpz = 0.0f;
    const int t1 = off[i + 1];
    for (int t = off[i]; t < t1; ++t) {
        const int j = idx[t];
        if (j == i) continue;
        const bool sand = j < n;
        const float4 pj = sand ? A[j] : solid[j - n];
        float ijx = pi.x - pj.x, ijy = pi.y - pj.y, ijz = pi.z - pj.z;
        if (sqrtf(dot3(ijx, ijy, ijz, ijx, ijy, ijz)) == 0.0f) {
            ijx = 0.0f;
            ijy = 0.00001f;
            ijz = 0.0f;
        }
        const float len = sqrtf(dot3(ijx, ijy, ijz, ijx, ijy, ijz));
        if (len > k.d) continue;
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
            const float4 xoj = xold[j];
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
    }
    float4 r;
    r.x = rclamp(pi.x + dpx, k.r, k.hi[0]);
    r.y = rclamp(pi.y + dpy, k.r, k.hi[1]);
    r.z = rclamp(pi.z + dpz, k.r, k.hi[2]);
    r.w = 0.0f;
    B[i] = r;
}

hipError_t launch_sand_substep(const float4* A, const float4* xold, const float4* solid,
                               const int* off, const int* idx, int n, SandK k, float4* B,
                               hipStream_t s) {
    hipLaunchKernelGGL(k_sand_substep, dim3(grid_for(n)), dim3(kBlock), 0, s, A, xold, solid, off,
                       idx, n, k, B);
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
    hipLaunchKernelGGL(k_sand_vupdate, dim3(grid_for(n)), dim3(kBlock), 0, s, A, xold_to_new, vout,
                       n, dt);
    return hipGetLastError();
}

}  // namespace lus
