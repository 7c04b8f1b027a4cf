// lus_internal.hpp -- types shared by the HIP kernels (lus_kernels.hip) and the
// host side of the C ABI (lus_capi.cpp).  Not installed; the public surface is
// include/lustrine_gpu.h.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "lustrine_gpu.h"

namespace lus {

// Per-kernel timing of the calling host thread.  While `on`, every kernel the
// library launches records an event pair stamped at that dispatch's start and
// end (start[q], stop[q]; phase[q] is the caller's label); `failed` is set if
// an event could not be created (that launch then goes untimed).
struct KernelTimes {
    bool on = false;
    bool failed = false;
    std::vector<hipEvent_t> start, stop;
    std::vector<int> phase;
};
KernelTimes& kernel_times();

// Derived constants of the PBF substep, all float32, computed once on the host
// (pbf_constants in lus_capi.cpp).  The kernels evaluate
//   W(r)  = P6 * t^3,          t = hf2 - f2 * r2            (r2 <= h2)
//   dW(r) = -s * r_vec,        s = SPf * g * g / r, g = hf - f * r
// and keep the accumulated t^3 unnormalised (c_rho = m * P6 / rho0).
// One scene of a batch: box clamp bounds (scene coordinates) and the y offset
// of its cells in the batch grid (scenes are stacked along y, one empty cell
// layer apart, so no stencil reaches another scene).
struct SceneBox {
    float lo[3], hi[3];
    int yoff;
    int pad;
};

struct PbfK {
    float h2, hf, hf2, f, f2, SPf, c_rho, mr0sq, eps, inv_tq3, scorr_k, neg_inv_r0, cW, cs;
    float lo[3], hi[3], g[3];
    int gX, gY, gZ, gXZ, n_scorr;
    // cell keys are ((y gX + x) nph + c) gZ + z: nph = 1 is get_cell_id's
    // order; the mixed scene splits each cell column into phase runs (nph = 3,
    // c = 0 fluid, 1 sand, 2 solid), so a stencil row is nph slot ranges
    int nph;
    int exact;                  // 1: oracle operation order; 0: fused fast path
    const SceneBox* scenes;     // batch table, nullptr for a single scene
    const int* sid;             // scene of each particle (input order), batch only
    int* sid_out;               // scene of each particle in cell order, batch only
    // mixed sand + fluid scene (nullptr ph: every particle is fluid)
    const int* ph;              // phase (0 fluid, 1 sand, 2 solid), input order
    int* ph_out;                // phase in cell order
    float mix_d, mix_collision, mix_friction, mix_mu_s, mix_mu_k;
    // a float |ij|^2 above it has sqrtf(|ij|^2) > mix_d: (d (1 + 2^-19))^2 or more
    float mix_d2_far;
    // fused mode: spiky's SPf and s_corr's 1 / t(dq)^3 moved out of the
    // neighbour loops: -scorr_k / t(dq)^12, SPf * neg_inv_r0, mr0sq * SPf^2
    float f_scorr12, f_spf_neg_inv_r0, f_mr0sq_spf2;
};

// simulate_sand constants (src/Simulate.cpp:159-163 and the Simulation fields).
struct SandK {
    float r, d, h2, w, cs;
    float mass_n, mass_sum;   // neighbor_mass and mass + neighbor_mass
    float hi[3], g[3], player[3];
    float collision, friction, mu_s, mu_k;
    float attract_radius, blow_radius, attract_coeff, blow_coeff;
    int gX, gY, gZ, gXZ;
    int attract, blow, prev_attract;
    int sym_limit;            // slots below it receive symmetric pushes
    // 0 simulate_sand, 1 simulate_sand_credits, 2 simulate_sand_v3
    // (src/Simulate.cpp:156-325, 327-510, 513-667)
    int variant;
    float bcoef;              // boundary_collision_coeff of simulate_sand_v3
    float dom[3];             // domainX/Y/Z: simulate_sand_v3's walls
    float d2_far;             // a squared length above it has sqrtf > d: never a contact
};

// Records a failure for lus_gpu_last_error (lus_capi.cpp) and returns `code`.
int report_error(int code, const char* fmt, ...);

// The simulate_fluid state of a context (lus_fluid.hip) and the context's
// device, stream and fluid slot for the translation units outside lus_capi.cpp.
struct FluidState;
void fluid_free(FluidState* f);
int ctx_device(const LusGpu* c);
hipStream_t ctx_stream(const LusGpu* c);
FluidState*& ctx_fluid(LusGpu* c);

// ---- launchers (lus_kernels.hip); all enqueue on `s` and return hipError_t ----
// prefix scan of m ints: out[k] = sum in[0..k); zero_in clears `in` after use.
// tile: the sort tile, ints per scan block (1024, 2048 or 4096); tile_sums
// needs scan_tile_count(m, tile) + 1 ints.
constexpr int kScanTileCells = 2048;
constexpr int kScanTileMin = 1024;
int scan_tile_count(int m, int tile = kScanTileCells);
bool scan_tile_valid(int tile);
hipError_t launch_exclusive_scan(const int* in, int* out, int m, int* tile_sums, int zero_in,
                                 hipStream_t s, int tile = kScanTileCells);
// The same scan in one pass (decoupled look-back).  `in` is left as it is
// (clear it afterwards: launch_scatter's clear); status and status_next: two
// rows of scan_lookback_rows(m) ints, both zero before the first call; the
// call reads status and zeroes status_next, so successive calls alternate
// the rows.  out must not alias in.
int scan_lookback_rows(int m);
// a[0..m) = 0 with a kernel (graph capture safe, unlike a memset node)
hipError_t launch_zero_ints(int* a, int m, hipStream_t s);
hipError_t launch_scan_lookback(const int* in, int* out, int m, int* status, int* status_next, int tile,
                                hipStream_t s);

hipError_t launch_cell_keys(const float4* pos, int n, LusGrid g, int* key, hipStream_t s);
hipError_t launch_count_keys(const int* key, int n, int* counts, int* slot, hipStream_t s);
// clear (nullable): clear_n ints zeroed as well (the histogram, for the next sort)
hipError_t launch_scatter(const int* key, const int* slot, const int* cell_start, int n, int* tmp,
                          hipStream_t s, int* clear = nullptr, int clear_n = 0);
// generic stable rank: perm[o] = i for the stable order; key_sorted[o] = key[i]
hipError_t launch_stable_rank(const int* tmp, const int* key, const int* cell_start, int n,
                              int* perm, int* key_sorted, hipStream_t s);
hipError_t launch_gather4(const float4* src, const int* perm, int n, float4* dst, hipStream_t s);
// dst[3 q .. 3 q + 2] = src[r].xyz, q = ids ? ids[r] : r (ids a permutation of 0..n-1)
hipError_t launch_pack3(const float4* src, const int* ids, int n, float* dst, hipStream_t s);
// *out += particles of the neighbour-list buffer nb whose header marks the
// re-test path (*out must be cleared by the caller)
hipError_t launch_pbf_count_retest(const unsigned long long* nb, int n, int* out, hipStream_t s);
hipError_t launch_gather1(const int* src, const int* perm, int n, int* dst, hipStream_t s);
hipError_t launch_stream_copy(const float4* src, float4* dst, size_t n4, hipStream_t s);

// fluid-path neighbor lists (parity entry point)
hipError_t launch_fluid_nbr(const float4* sand_sorted, const int* sand_key, const int* sand_cs,
                            const int* sand_orig, int n_sand, int sand_start,
                            const float4* solid_sorted, const int* solid_cs, const int* solid_ref,
                            LusGrid g, float h2, const int* inv_perm, const long* off, int* cnt,
                            int* idx, int fill, hipStream_t s);

// Neighbour sets of a step, one buffer of nbr_words(n) 64-bit words.  Flat
// passes: neighbour lists in ascending slot order.  First a header per slot,
// hdr[i] = int4(b0, b1, b2, w): b_L is the slot of the first neighbour in the
// stencil's y layer L (cell y - 1, y, y + 1; slot i when the layer has none),
// w = |N(i)| | n0 << 16 | n01 << 24 with n0 the neighbours in layer 0 and n01
// those in layers 0 and 1.  Then per run of 64 slots (w = i / 64, l = i % 64)
// kNbrGroups groups of 64 int4, each int4 eight 16-bit entries: entry t of
// slot i is in int4 number (w * kNbrGroups + t / 8) * 64 + l, bits
// 16 (t % 8) .. 16 (t % 8) + 15 of the 128, and holds j - b_L for neighbour j
// of layer L.  A particle with more than kNbrCap neighbours, or one whose
// neighbours lie more than 65535 slots past their layer's first, has
// w = kNbrRetest and re-tests its candidates in every pass.  Tiled passes: 9
// row masks per particle, masks[r * n + i].  A step uses one of the two forms.
constexpr int kNbrCap = 128;  // about 4x the rest-state count (a K = 1 solve compresses 2.5x)
constexpr int kNbrGroups = kNbrCap / 8;
constexpr int kNbrRetest = 0xFFFF;
// particles of one PBF context: float4 data of any slot lies within 4 GiB
constexpr int kMaxParticles = (1 << 28) - 1;
inline size_t nbr_words(size_t n) {
    const size_t int4s = n + (n + 63) / 64 * (size_t)kNbrGroups * 64;
    return 2 * int4s > 9 * n ? 2 * int4s : 9 * n;
}

// Tiled neighbour passes (lus_pbf.hip): tiles of at most kTileT slots of one
// (y, x) cell column, built per step from cells in chunks of zt <= kTileZMax
// along z; each block stages its tile's 9 stencil rows in LDS.
constexpr int kTileT = 128;
constexpr int kTileZMax = 32;
struct TileArgs {
    const int4* tiles = nullptr;  // nullptr: one thread per slot, global loads
    const int* ntiles = nullptr;  // device tile count
    int grid = 0;                 // blocks of kTileT threads, a multiple of 8
    int lcap = 0;                 // staged float4 per staged array
    int block = kBlock;           // threads per block of a flat launch (64, 128 or 256)
    int stage = 1;                // flat list build: wave-staged candidates (0: per-lane gathers)
    int sub = 1;                  // lanes per particle of the list walks (small fused scenes: kSubLanes)
};
// A fused plain scene of at most kSubMaxN particles walks its lists with
// kSubLanes lanes per particle (lus_pbf.hip, "Small scenes").
constexpr int kSubLanes = 4;
constexpr int kSubMaxN = 131072;
int tile_chunks(const PbfK& k, int zt);
// tcount: tile_chunks + 1 ints (last one 0); toff: the same; scan_tiles:
// scan_tile_count(tile_chunks + 1) + 1; tiles: n / kTileT + tile_chunks int4.
// *ntiles = toff[tile_chunks].
hipError_t launch_tiles(const int* cs, const int* key2, PbfK k, int zt, int* tcount, int* toff, int* scan_tiles,
                        int4* tiles, hipStream_t s);

// PBF substep.  `gate` (nullable): a device int; the kernel does nothing when
// it holds 0 (the conditional plain rerun of the learned step).  `ta`: tiled
// launch (ta.tiles non-null) or one thread per slot.
hipError_t launch_pbf_key(const float4* x, const float4* v, int n, PbfK k, float dt, int* key,
                          int* slot, int* counts, hipStream_t s);
hipError_t launch_pbf_permute(const int* tmp, const int* key, const int* cell_start,
                              const float4* x, const float4* v, const int* id, int n, PbfK k,
                              float dt, float4* x2, float4* xp, int* id2, int* key2,
                              const int* gate, hipStream_t s);
// x2 only (old positions in cell order), as launch_pbf_permute writes it
hipError_t launch_pbf_restore_old(const int* tmp, const int* key, const int* cell_start,
                                  const float4* x, int n, float4* x2, const int* gate, hipStream_t s);
// masks: the neighbour-set buffer (nbr_words(n) words); build = 1 on the first
// lambda pass of a step
hipError_t launch_pbf_lambda(const float4* xa, int* key2, const int* cs, int n, PbfK k, float4* xl,
                             int* maxc, unsigned long long* masks, const float4* xp, int build,
                             const int* gate, const TileArgs& ta, hipStream_t s);
hipError_t launch_pbf_deltap(const float4* xl, int* key2, const int* cs, int n, PbfK k, float4* xa,
                             const float4* xold, float4* vout, float dt, unsigned long long* masks,
                             const float4* xp, const int* gate, const TileArgs& ta, hipStream_t s);
hipError_t launch_pbf_velocity(const float4* xa, const float4* xold, int n, float dt,
                               float4* vout, hipStream_t s);
// XSPH + commit; maxd (nullable, 2 * blocks ints): per-block max of the
// density constraint C_i = rho_i / rho0 - 1 at the final positions
// (ordered() ints), then per-block sums of max(0, C_i) (float bits)
// The next step's predicted cell keys, computed by the XSPH pass from the
// positions and velocities it commits (launch_pbf_key's outputs: key, slot
// and the histogram counts, which must be zero), with the next step's dt.
struct NextKeys {
    int* key = nullptr;  // nullptr: none
    int* slot = nullptr;
    int* counts = nullptr;
    float dt = 0.0f;
};
// nk (nullable): also write the next step's keys (one-thread and four-lane
// passes; not with tiles)
hipError_t launch_pbf_xsph(const float4* xa, const float4* vt, int* key2, const int* cs, int n,
                           PbfK k, float4* vout, float4* xout, unsigned long long* masks,
                           const float4* xp, int* maxd, const int* gate, const TileArgs& ta, hipStream_t s,
                           const NextKeys* nk = nullptr);

// ---- mixed sand + fluid passes (k.ph, k.ph_out set) ----
hipError_t launch_mix_lambda(const float4* xa, int* key2, const int* cs, int n, PbfK k, float4* xl,
                             unsigned long long* masks, const float4* xp, int build, int stage, hipStream_t s);
hipError_t launch_mix_correct(const float4* xl, int* key2, const int* cs, int n, PbfK k, float4* xa,
                              const float4* xold, float4* vout, float dt, unsigned long long* masks,
                              const float4* xp, hipStream_t s);
hipError_t launch_mix_xsph(const float4* xa, const float4* vt, int* key2, const int* cs, int n, PbfK k,
                           float4* vout, float4* xout, unsigned long long* masks, const float4* xp, int* maxd,
                           hipStream_t s);

// ---- learned warm start ----
// Per-particle features of the predicted positions (build pass: also writes
// the neighbour masks, like the first lambda pass).  Exactly one of
//   fbf:  bf16 [n][32], (raw - mu) * inv_sigma       (mu/inv_sigma: 32 floats each)
//   frec: fp32 [rows][32] raw, only slots p with (p + rec_off) % rec_stride == 0,
//         row (p - p0) / rec_stride, p0 the first such slot
// is non-null.
constexpr int kFeat = 32;
hipError_t launch_pbf_features(int* key2, const int* cs, int n, PbfK k, unsigned long long* masks,
                               const float4* xp, const float4* xold, const float* mu,
                               const float* inv_sigma, void* fbf, float* frec, int rec_stride,
                               int rec_off, const TileArgs& ta, hipStream_t s);
// training target of the recorded slots: tgt[row] = xa[p] - xp[p]
hipError_t launch_pbf_record_target(const float4* xa, const float4* xp, int n, int rec_stride,
                                    int rec_off, float* tgt, hipStream_t s);
// reduce the XSPH block maxima and sums (maxd: 2 * nblk): e = max(0, max C),
// hist[step % cap] = e, hist[cap + step % cap] = mean_i max(0, C_i);
// flag = learned && !(e <= bound); counters[0] += 1, counters[1] += flag
hipError_t launch_pbf_decide(const int* maxd, int nblk, int n, float bound, int learned, int* flag,
                             int* counters, float* hist, int cap, hipStream_t s);

// MLP 32 -> 64 (SiLU) -> 64 (SiLU) -> 3, parameters (fp32) in this order:
constexpr int kMlpIn = 32, kMlpHid = 64, kMlpOut = 3;
constexpr int kMlpW1 = 0;                              // [32][64]
constexpr int kMlpB1 = kMlpW1 + kMlpIn * kMlpHid;      // [64]
constexpr int kMlpW2 = kMlpB1 + kMlpHid;               // [64][64]
constexpr int kMlpB2 = kMlpW2 + kMlpHid * kMlpHid;     // [64]
constexpr int kMlpW3 = kMlpB2 + kMlpHid;               // [64][3]
constexpr int kMlpB3 = kMlpW3 + kMlpHid * kMlpOut;     // [3]
constexpr int kMlpParams = kMlpB3 + kMlpOut;           // 6467
// bf16 B-operand images of W1, W2, W3 (bytes)
constexpr int kMlpPackBytes = (4 + 8 + 2) * 64 * 16;

hipError_t launch_mlp_pack(const float* params, void* packed, hipStream_t s);
// out (nullable): [n][3] = out_scale * mlp(X); xa (nullable): clamp(xp + that)
hipError_t launch_mlp_forward(const void* X, int n, const float* params, const void* packed,
                              float out_scale, float* out, const float4* xp, float4* xa,
                              const PbfK& k, hipStream_t s);
// training kernels (row-major fp32 / bf16 matrices)
hipError_t launch_gemm_bf16(const void* A, long sam, long sak, const void* B, long sbk, long sbn,
                            float* C, int ldc, int M, int N, int K, int kchunk, int ones_row,
                            hipStream_t s);
hipError_t launch_sum_partials(const float* part, int chunks, int MN, float* out, hipStream_t s);
hipError_t launch_bias_silu(float* z, const float* bias, int rows, int N, void* h, hipStream_t s);
hipError_t launch_out_loss(const float* z3, const float* b3, const float* y, int rows,
                           float inv_scale, void* dbf, float* err, hipStream_t s);
hipError_t launch_silu_back(const float* dh, const float* z, long tot, void* dzb, hipStream_t s);
hipError_t launch_to_bf16(const float* a, long tot, void* b, hipStream_t s);
hipError_t launch_normalize_rows(const float* x, const int* rows, long nrows, const float* mu,
                                 const float* inv_sigma, void* xb, hipStream_t s);
hipError_t launch_gather_rows(const float* y, const int* rows, long nrows, int w, float* out,
                              hipStream_t s);
hipError_t launch_sum_into(const float* v, int n, float* acc, hipStream_t s);
hipError_t launch_col_stats(const float* x, long rows, int w, double* out, hipStream_t s);
hipError_t launch_adam(float* p, const float* g, float* m, float* v, int np, float lr, float b1,
                       float b2, float eps, float c1, float c2, hipStream_t s);

// sand (simulate_sand)
hipError_t launch_sand_predict_key(float4* x, float4* v, int* att, int n, SandK k, float dt, int* err,
                                   float4* xs, int* key, int* slot, int* counts, hipStream_t s);
hipError_t launch_sand_permute(const int* tmp, const int* key, const int* cell_start,
                               const float4* x, const float4* xs, const float4* v, const int* att,
                               int n, float4* x2, float4* xs2, float4* v2, int* att2, int* perm,
                               int* key2, hipStream_t s);
// fill with fail non-null: idx holds cap entries; a total past cap (or a
// negative one) writes nothing and sets *fail, and a substep with *fail set
// does nothing
hipError_t launch_sand_nbr(const float4* xs, const int* key2, const int* sand_cs, int n,
                           const float4* solid, const int* solid_cs, SandK k, const int* off,
                           int* cnt, int* cnt1, int* idx, int fill, hipStream_t s, long cap = 0,
                           int* fail = nullptr);
hipError_t launch_sand_substep(const float4* A, const float4* xold, const float4* solid,
                               const int* off, const int* idx, int n, SandK k, float4* B, int* att,
                               hipStream_t s, const int* fail = nullptr);
hipError_t launch_sand_vupdate(const float4* A, float4* xold_to_new, float4* vout, int n, float dt,
                               hipStream_t s);

// grid-state consumers (lus_kernels.hip).  Cell box lo..hi (x, y, z; inclusive)
// of a grid with cell starts cs (sand slots) and ocs (static solids, reference
// slots solid_ref); *out += the listed ids (all_ids) or those below sand_end.
// *out must be zeroed first; the grid is gX x . x gZ, y-major.
hipError_t launch_cell_box_count(const int* cs, const int* ocs, const int* solid_ref, const int* lo,
                                 const int* hi, int gX, int gZ, int sand_end, int all_ids, int* out,
                                 hipStream_t s);
// out[k] = (cs[cells[k]], cs[cells[k] + 1]); cells must be valid cell ids
hipError_t launch_cell_ranges(const int* cs, const int* cells, int n, int2* out, hipStream_t s);
// first k slots (plain slot order) with |x - c| <= h on every axis:
// block_counts[b] = matches of slots [256 b, 256 b + 256), (n + 255) / 256
// entries; after an exclusive scan into block_off, box_write stores the slots
// and positions of the matches of rank < k
hipError_t launch_box_count(const float4* x, int n, const float* c, const float* h, int* block_counts,
                            hipStream_t s);
hipError_t launch_box_write(const float4* x, int n, const float* c, const float* h, const int* block_off, int k,
                            int* slots, float4* pos, hipStream_t s);
// scene construction (init_chunk_from_grid): flags[t] of the F = X Y Z s^3
// fine lattice points; after an exclusive scan into rank, chunk_write stores
// the positions (and per-cell colours, col nullable) of the kept points
hipError_t launch_chunk_flags(const int* cells, int X, int Y, int Z, int sub, int mask_out, long F, int* flags,
                              hipStream_t s);
hipError_t launch_chunk_write(const int* cells, const float4* cell_colors, int X, int Y, int Z, int sub,
                              int mask_out, long F, const float* origin, float cs, const int* rank, float4* pos,
                              float4* col, hipStream_t s);
// voxel model -> grid cells and colours; *occupied += set voxels
hipError_t launch_grid_from_voxels(const unsigned char* vox, int sx, int sy, int sz, const uchar4* palette,
                                   int* cells, float4* colors, int* occupied, hipStream_t s);
// a[mv[k].x] = a_old[mv[k].y] for the four arrays; t*: m staging entries each
hipError_t launch_slot_moves(const int2* mv, int m, float4* a0, float4* a1, float4* a2, int* a3,
                             float4* t0, float4* t1, float4* t2, int* t3, hipStream_t s);

}  // namespace lus
