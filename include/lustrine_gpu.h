/*
 * lustrine_gpu.h -- C ABI of the MI355X (gfx950) particle substep.
 *
 * Plain C: pointers, sizes and POD structs only.  Every entry point returns 0
 * on success and a nonzero LUS_E* code on failure; the message of the last
 * failure is available from lus_gpu_last_error().  Nothing here falls back to
 * the CPU: a HIP failure is reported, the step is not computed.
 *
 * Host arrays are AoS float[3] per particle (the layout of glm::vec3 arrays in
 * the reference's Lustrine::Simulation, src/Simulation.hpp:173-227) unless
 * stated otherwise.
 */
#ifndef LUSTRINE_GPU_H
#define LUSTRINE_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LUS_OK 0
#define LUS_EINVAL 1   /* bad argument or size */
#define LUS_EHIP 2     /* a HIP runtime call failed */
#define LUS_ESTATE 3   /* call out of order (e.g. step before init) */
#define LUS_ENOMEM 4

/* Uniform grid of the reference (src/Lustrine.cpp:546-558,
 * src/neighbors/Utils.hpp:24-33): cell = (int)(y/cs)*gX*gZ + (int)(x/cs)*gZ +
 * (int)(z/cs), y-major, no clamp. */
typedef struct LusGrid {
    float cell_size;
    int gridX, gridY, gridZ;
} LusGrid;

/* Position-based-fluids substep constants.  Derived constants (poly6 and spiky
 * coefficients, 1/rho0, ...) are computed by the library from these. */
typedef struct LusPbfParams {
    float h;             /* kernel radius (Simulation::kernelRadius) */
    float f;             /* Simulation::kernelFactor */
    float cell_size;
    int gridX, gridY, gridZ;
    float lo[3], hi[3];  /* box clamp [r, D - r] per axis */
    float gravity[3];
    float mass, rest_density, epsilon;   /* Simulation::mass, rest_density, relaxation_epsilon */
    float scorr_k, scorr_dq;             /* Simulation::s_corr_k, s_corr_dq */
    int scorr_n;                         /* Simulation::s_corr_n (integer power) */
    float c_xsph;                        /* Simulation::c_xsph */
    int iterations;                      /* Jacobi solver iterations */
    /* 1: every float operation in the order of the CPU oracle, IEEE divide and
     *    sqrt, no fused multiply-add (results bit-identical to the oracle);
     * 0: fused multiply-add and hardware reciprocal square root in the
     *    neighbour loops, and h^2 - r^2 by fused multiply-adds in the
     *    neighbour test (results within a few ulp of mode 1; the two tests
     *    differ only for pairs within rounding of r = h). */
    int exact;
} LusPbfParams;

/* simulate_sand constants (src/Simulate.cpp:159-163 locals and the
 * Simulation fields the function reads). */
typedef struct LusSandParams {
    float particleRadius, particleDiameter, kernelRadius, mass;
    float domain[3];                     /* domainX, domainY, domainZ */
    LusGrid grid;
    float gravity[3];
    float collision_coeff, friction_coeff, mu_s, mu_k;
    int substeps;
    float attract_radius, blow_radius, attract_coeff, blow_coeff;
    /* Simulation::ptr_solid_start relative to the first sand slot: the
     * symmetric push of find_neighbors_uniform_grid_v1 only adds entries to
     * slots below it (src/neighbors/Neighbors.cpp:351-352).  Any value
     * >= n_sand means every sand slot. */
    int ptr_solid_start;
} LusSandParams;

/* ------------------------------------------------------------------ */
/* process-wide                                                        */
/* ------------------------------------------------------------------ */
/* Copies the message of the last failure (NUL-terminated, truncated to n).
 * Returns the code of the last failure (0 if none).  No reference
 * counterpart: the reference reports nothing (void + assert). */
int lus_gpu_last_error(char* buf, int n);
/* Number of visible devices; 0 when no GPU (does not fail). */
int lus_gpu_device_count(void);

/* ------------------------------------------------------------------ */
/* building blocks (host in, host out; the library stages device memory) */
/* ------------------------------------------------------------------ */
/* get_cell_id over n points (src/neighbors/Utils.hpp:24-33). */
int lus_cell_ids(const LusGrid* g, int n, const float* pos, int* out);

/* Sorting::counting_sort (src/neighbors/Sorting.cpp:11-33, declared at
 * src/neighbors/Sorting.hpp:13-19): sorted[k] = index of the k-th element in
 * stable ascending key order; counts[0..num_cells] receives the table the
 * reference leaves behind (first slot of each key; counts[num_cells] = n). */
int lus_counting_sort(int* counts, const int* keys, size_t n, size_t num_cells, int* sorted);

/* The same counting sort on device buffers, allocating nothing.  No counterpart
 * in the reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops, harnesses that forbid hipMalloc).
 * keys[n], counts[num_cells + 1] and sorted[n] are device pointers on the
 * current device; the work is enqueued on `stream` (a hipStream_t, NULL for
 * the null stream) and the call returns without waiting.  Every intermediate
 * lives in one static device array per GPU (about 41 MB), sized for at most
 * LUS_NOALLOC_MAX_KEYS keys and LUS_NOALLOC_MAX_CELLS cells; larger sizes are
 * rejected with LUS_EINVAL.  All calls on one GPU share that array: calls
 * that overlap on one GPU (other threads, other streams) race. */
#define LUS_NOALLOC_MAX_KEYS (1 << 21)
#define LUS_NOALLOC_MAX_CELLS (1 << 22)
int lus_counting_sort_noalloc(const int* keys, int n, int num_cells, int* counts, int* sorted, void* stream);

/* find_neighbors_uniform_grid (src/neighbors/Neighbors.cpp:366-450): neighbor
 * lists of the n_sand moving particles in the reference's order.  Indices in
 * idx are reference slots: sand k -> sand_start + k, solid k -> solid_index[k].
 * off has n_sand+1 entries.  If cap is too small, *total receives the needed
 * count, nothing is written and LUS_EINVAL is returned. */
int lus_neighbors_fluid(const LusGrid* g, float kernel_radius,
                        int n_sand, const float* sand_pos, int sand_start,
                        int n_solid, const float* solid_pos, const int* solid_index,
                        long* off, int* idx, long cap, long* total);

/* ------------------------------------------------------------------ */
/* device context                                                      */
/* ------------------------------------------------------------------ */
typedef struct LusGpu LusGpu;

/* Measured HBM ceiling: copies `bytes` (multiple of 16) between two device
 * buffers `reps` times with a float4 copy kernel (k_stream_copy) on the
 * context stream; *us = average microseconds per copy.  Read + write traffic
 * per copy is 2 * bytes.  No reference counterpart (measurement support). */
int  lus_gpu_stream_copy(LusGpu* ctx, size_t bytes, int reps, float* us);
int  lus_gpu_create(LusGpu** out, int device);
void lus_gpu_destroy(LusGpu* ctx);
int  lus_gpu_sync(LusGpu* ctx);

/* --- plain PBF scene (fluid particles in a closed box) --- */
/* Uploads n particles (host AoS float[3]).  vel may be NULL (zero).
 * Particle identities 0..n-1 are carried through the per-step reordering. */
int lus_pbf_init(LusGpu* ctx, const LusPbfParams* p, int n, const float* pos, const float* vel);
/* Many independent scenes in one set of launches.  Scene s has counts[s]
 * particles (concatenated in pos/vel, scene coordinates) in the box
 * [0, boxes[3s..3s+2]]; its clamp box is [p->lo, box - p->lo].  p gives the
 * shared solver constants (gridX/Y/Z and hi are ignored).  The scene grids are
 * stacked along y one empty cell layer apart in one batch grid, so each scene
 * computes exactly what it computes alone.  Particle ids run over the
 * concatenation; lus_pbf_download with order = 1 returns that order. */
int lus_pbf_batch_init(LusGpu* ctx, const LusPbfParams* p, int nscenes, const int* counts,
                       const float* boxes, const float* pos, const float* vel);
/* Mixed sand + fluid over static solid particles: one particle array with
 * phase[i] = 0 fluid, 1 sand, 2 solid (solids never move).  Fluid follows the
 * PBF step with sand and solids counted in its density (their lambda is 0);
 * sand follows the contact + friction projection of simulate_sand
 * (src/Simulate.cpp:226-311) against sand (half weight, friction), fluid
 * (half weight, contact only) and solid (full weight, friction) neighbours
 * within `diameter`; p->iterations Jacobi iterations for both; XSPH over
 * fluid pairs.  Steps with lus_pbf_step; lus_pbf_download returns every
 * phase (the ids give the input order).  No reference counterpart for the
 * coupling (the reference runs sand or fluid, src/LustrineWrapper.cpp:666-677). */
typedef struct LusMixParams {
    float diameter;                          /* contact distance d = 2r */
    float collision, friction, mu_s, mu_k;   /* roles of src/Simulate.cpp:159-163 */
} LusMixParams;
int lus_mix_init(LusGpu* ctx, const LusPbfParams* p, const LusMixParams* mp, int n, const float* pos,
                 const float* vel, const int* phase);
/* Enqueues `steps` substeps of length dt on the context stream. */
int lus_pbf_step(LusGpu* ctx, float dt, int steps);
/* on = 1: lus_pbf_step replays a hipGraph holding two captured substeps
 * (captured on first use and again when dt changes); on = 0: plain launches. */
int lus_pbf_set_graph(LusGpu* ctx, int on);
/* Neighbour passes: tiled = 1 runs one block per tile of at most 128
 * consecutive slots of one (y, x) cell column, cells cut into chunks of zt
 * (1..32) cells along z, with the tile's 27-cell neighbourhood staged in LDS;
 * tiled = 0 (default, faster on gfx950) runs one thread per slot with global
 * loads.  Same results bit for bit.  No reference counterpart (launch
 * configuration). */
int lus_pbf_set_tiling(LusGpu* ctx, int tiled, int zt);
/* List build of the flat passes (the first lambda pass of a step): staged = 1
 * (default) loads each wave's candidate ranges of a stencil row once into LDS
 * and tests them there; 0 gathers every candidate per lane from global memory.
 * Same results bit for bit.  No reference counterpart (launch configuration;
 * the scan replaced is src/neighbors/Neighbors.cpp:306-361). */
int lus_pbf_set_build(LusGpu* ctx, int staged);
/* Lanes per particle in the list walks (lambda, delta-p, XSPH) of fused plain
 * steps: 0 (default) = 4 for scenes of at most 131,072 particles, else 1;
 * 1 or 4 forces it.  With 4 lanes each walks every fourth group of eight
 * entries and the partial sums are combined in a fixed order: deterministic,
 * not in the oracle's summation order (fused mode only; exact mode always
 * walks with one lane).  No reference counterpart (launch configuration). */
int lus_pbf_set_lanes(LusGpu* ctx, int lanes);
/* Sort tile of the step's counting sort: cells per block of its prefix scan
 * over the cell histogram, 1024, 2048 or 4096; 0 (default): 4096 for grids of
 * at least 2^19 cells, else 2048.  Same sorted order.
 * Kept across scene inits; chosen per size bucket by the autotuner.  The scan
 * replaced is the inclusive prefix of Sorting::counting_sort
 * (src/neighbors/Sorting.cpp:11-33). */
int lus_pbf_set_sort_tile(LusGpu* ctx, int cells);
/* on = 1 (default): the XSPH pass of a plain step also predicts the next
 * step's positions and writes their cell keys and histogram slots, and the
 * next step skips its key kernel when it runs with the same dt (any other dt,
 * a learned or mixed step, tiles, or a new scene run the key kernel).  The
 * same keys, so the same results bit for bit; phase 0 then has no launch and
 * its work is timed in phase 6.  0: every step runs the key kernel.  No
 * reference counterpart (kernel fusion of predict (src/Simulate.cpp:48-51)
 * and get_cell_id into the previous step's commit). */
int lus_pbf_set_key_fusion(LusGpu* ctx, int on);
/* Threads per block (64, 128 or 256; default 256) of the per-slot neighbour
 * passes.  Same results.  No reference counterpart (launch configuration). */
int lus_pbf_set_launch(LusGpu* ctx, int block);
/* Runs `steps` substeps with a hipEvent between steps (between step pairs in
 * graph mode, each step of a pair getting half); ms[k] = step k's time. */
int lus_pbf_step_latency(LusGpu* ctx, float dt, int steps, float* ms);
/* Downloads the state.  order = 0: internal cell order (ids[] says which
 * particle each row is); order = 1: original particle order.  Any output
 * pointer may be NULL. */
int lus_pbf_download(LusGpu* ctx, int order, float* pos, float* vel, int* ids);
/* Max over particles of the density constraint C_i = rho_i/rho0 - 1 measured
 * in the last lambda pass of the last step. */
int lus_pbf_last_max_constraint(LusGpu* ctx, float* out);
/* Runs warmup + steps substeps; *ms = elapsed of the `steps` (hipEvents on the
 * context stream, one before and one after).  When phase_ms is non-NULL
 * (array of LUS_PBF_NPHASES) the `steps` instead run with plain launches and
 * every kernel timed by an event pair stamped at its own dispatch's start and
 * end (hipExtLaunchKernelGGL; the gaps between kernels are not counted):
 * phase_ms accumulates per-phase kernel time, phase_launches the number of
 * times each phase ran, and *ms the sum of all kernel times. */
#define LUS_PBF_NPHASES 12
/* phases: 0 predict+cell key, 1 cell scan, 2 scatter, 3 stable permute,
 *         4 lambda pass (iterations 2..K), 5 delta-p pass, 6 XSPH + commit,
 *         7 first lambda pass of the step (builds the neighbour sets),
 *         8 warm-start features (build the neighbour sets), 9 warm-start
 *         network, 10 density decision (learned mode or density check),
 *         11 tile list of the tiled neighbour passes (phase 0 has no launch
 *         in a step whose keys the previous step's XSPH wrote:
 *         lus_pbf_set_key_fusion) */
int lus_pbf_bench(LusGpu* ctx, float dt, int warmup, int steps, float* ms,
                  float* phase_ms, int* phase_launches);
/* Average device time per launch (microseconds) of the neighbour-pass kernels
 * on the current state: us[0] lambda pass, us[1] delta-p pass, us[2] XSPH
 * pass, us[3] the first lambda pass of a step (which also builds the
 * neighbour sets), each launched `reps` times back to back between two
 * hipEvents on the context stream.  Outputs go to scratch buffers; the scene
 * state is left unchanged.  Requires at least one step since init. */
int lus_pbf_time_kernels(LusGpu* ctx, int reps, float* us);

/* --- learned warm start of the PBF solve ---
 * No reference counterpart: the reference solves the density constraint with
 * one uncorrected pass (src/Simulate.cpp:83-154).  A per-particle network
 * maps 32 features of the predicted positions to a position correction; a
 * learned step applies it, runs `learned_iterations` plain Jacobi
 * iterations, measures e = max_i max(0, rho_i/rho0 - 1) at the final
 * positions (in the XSPH pass) and, when e > bound (or e is NaN), reruns the
 * plain step (LusPbfParams::iterations) from the predicted positions.  The
 * decision stays on the device (no host synchronisation); a rerun gives the
 * plain step's results bit for bit.
 *
 * Features (per particle i, over its neighbour set N(i) at the predicted
 * positions x*, offsets d = x*_j - x*_i, r = |d|):
 *   [0]         rho_i/rho0 - 1 (self included)
 *   [1..3]      (x*_i - x_i)/h
 *   [4+3b..6+3b] b = 0..7: mean d/h over j != i with min(7, (int)(8 r/h)) = b
 *   [28]        |N(i)| - 1 over 32
 *   [29], [30]  mean r/h and mean r^2/h^2 over j != i
 *   [31]        min(1, distance to the nearest clamp wall / h)
 * Network: 32 -> 64 (SiLU) -> 64 (SiLU) -> 3, bf16 weights and activations,
 * fp32 accumulation, output * out_scale = position correction.
 * Network file (LUS_NET_FLOATS floats): feature mean [32], feature 1/std [32],
 * parameters [LUS_MLP_PARAMS] = W1[32][64] b1[64] W2[64][64] b2[64] W3[64][3]
 * b3[3] (W[k][n]: input k -> output n), out_scale [1]. */
#define LUS_FEATURES 32
#define LUS_MLP_PARAMS 6467
#define LUS_NET_MU 0
#define LUS_NET_INV_SIGMA 32
#define LUS_NET_PARAMS 64
#define LUS_NET_OUT_SCALE (LUS_NET_PARAMS + LUS_MLP_PARAMS)
#define LUS_NET_FLOATS (LUS_NET_OUT_SCALE + 1)
#define LUS_DENSITY_HISTORY 4096

/* Uploads the network the learned step uses (kept across scene re-inits). */
int lus_pbf_set_network(LusGpu* ctx, const float* net, int nfloats);
/* learned = 1: learned steps (needs a network, LusPbfParams::iterations >= 1);
 * 0: plain steps.  check = 1 in plain mode: the XSPH pass also measures e and
 * records it (no rerun).  Resets the statistics; drops a captured graph. */
int lus_pbf_set_mode(LusGpu* ctx, int learned, int learned_iterations, float bound, int check);
/* Steps measured since the last reset, reruns among them, and for the last
 * min(steps, LUS_DENSITY_HISTORY, cap) steps (oldest first) e in errs and the
 * mean over particles of max(0, rho_i/rho0 - 1) in errs_mean (either may be
 * NULL); *nerrs receives their count.  In learned mode a step that reran
 * records the rerun's (plain) solution, any other the learned one. */
int lus_pbf_learned_stats(LusGpu* ctx, long long* steps, long long* reruns, float* errs,
                          float* errs_mean, int cap, int* nerrs);
/* Particles of the last step whose neighbour list could not be stored (more
 * than 128 entries, or a neighbour more than 65,535 slots past its stencil
 * layer's first) and that therefore re-tested their 27-cell candidates in
 * every pass of that step (0 before a step and for the tiled passes).  No
 * reference counterpart (the reference's lists are std::vectors). */
int lus_pbf_retest_count(LusGpu* ctx, long long* count);

/* Training of the network on the device.  A trainer holds a dataset of
 * (raw features, target) rows; the target of a recorded particle is the plain
 * solver's position correction x_K - x* of that step. */
typedef struct LusMlpTrainer LusMlpTrainer;
typedef struct LusMlpFitParams {
    int epochs, batch;           /* batch: rows per Adam step (multiple of 32) */
    float lr, lr_final;          /* cosine schedule from lr to lr_final */
    float val_fraction;          /* rows held out for validation (random) */
    unsigned long long seed;     /* initialisation and shuffling */
} LusMlpFitParams;
int  lus_mlp_trainer_create(LusGpu* ctx, long capacity, LusMlpTrainer** out);
void lus_mlp_trainer_destroy(LusMlpTrainer* t);
/* Appends host rows: feat [rows][32], target [rows][3]. */
int  lus_mlp_trainer_add(LusMlpTrainer* t, const float* feat, const float* target, long rows);
int  lus_mlp_trainer_rows(LusMlpTrainer* t, long* rows);
int  lus_mlp_trainer_download(LusMlpTrainer* t, long first, long rows, float* feat, float* target);
/* One plain step of the context's scene that also appends the features and
 * targets of the particles in cell-order slots p with (p + k) % stride == 0,
 * k = the number of earlier record calls on this trainer.  *added (nullable)
 * receives the number of rows appended. */
int  lus_pbf_record_step(LusGpu* ctx, LusMlpTrainer* t, float dt, int stride, long* added);
/* Fits a fresh network to the dataset (normalisation from the dataset,
 * LeCun-uniform initialisation, Adam with b1 0.9, b2 0.999, eps 1e-8, MSE of
 * target / out_scale with out_scale the RMS target component).  Mean losses
 * per epoch go to train_loss / val_loss (nullable, `epochs` floats each). */
int  lus_mlp_fit(LusMlpTrainer* t, const LusMlpFitParams* fp, float* train_loss, float* val_loss);
int  lus_mlp_trainer_network(LusMlpTrainer* t, float* net);
/* Loss and gradient of the training kernels on host data: X [rows][32]
 * (network inputs, rounded to bf16), Y [rows][3]; loss = mean over rows of
 * sum_c (o_c - Y_c * inv_scale)^2; grad [LUS_MLP_PARAMS]. */
int  lus_mlp_loss_grad(LusGpu* ctx, const float* params, int rows, const float* X, const float* Y,
                       float inv_scale, float* loss, float* grad);
/* The learned step's inference kernel on host features: out [rows][3] =
 * out_scale * network((feat - mu) * inv_sigma). */
int  lus_mlp_infer(LusGpu* ctx, const float* net, int rows, const float* feat, float* out);

/* --- sand scene: simulate_sand (src/Simulate.cpp:156-325) --- */
/* n_sand moving particles (pos, vel, attracted may be NULL for zero) and
 * n_solid static particles with their reference slots solid_index[k]. */
int lus_sand_init(LusGpu* ctx, const LusSandParams* p, int n_sand, const float* pos,
                  const float* vel, const int* attracted, int n_solid,
                  const float* solid_pos, const int* solid_index);
/* One simulate_sand call.  player = Bullet player position (3 floats, may be
 * NULL when both flags are 0).  By default the call waits once for the device
 * mid-step: the neighbour lists' total length sizes their buffer before they
 * are written (the lists have no fixed capacity, as the reference's vectors).
 * With a fixed capacity (lus_sand_set_list_capacity) it only enqueues. */
int lus_sand_step(LusGpu* ctx, float dt, int attract_flag, int blow_flag, const float* player);
/* entries > 0: the neighbour lists get a fixed buffer of that many entries
 * and lus_sand_step no longer waits for the device (the reference's lists
 * grow instead, src/neighbors/Neighbors.cpp:306-361; ~33 entries per
 * particle on the 1 M sand lattice).  A step whose lists would not fit writes
 * none, skips its substeps and leaves an invalid state; the next
 * lus_gpu_sync, lus_sand_download or lus_sand_neighbors reports it
 * (LUS_ESTATE) and clears the mark.  In the same way a simulate_sand_v3
 * prediction outside the grid is reported there (LUS_EINVAL).  0 (default):
 * the waiting step.  Reset by lus_sand_init.  No reference counterpart. */
int lus_sand_set_list_capacity(LusGpu* ctx, long entries);
/* on = 1: each lus_sand_step replays a hipGraph of the whole step, captured
 * once per launch configuration (dt, flags, player position when a flag is
 * set, the previous attract flag, variant, and the alternating buffers: a run
 * with fixed arguments uses two graphs); needs a fixed list capacity.  Same
 * results bit for bit.  Reset by lus_sand_init.  No reference counterpart. */
int lus_sand_set_graph(LusGpu* ctx, int on);
/* Which of the reference's sand Simulate_funs lus_sand_step computes (after
 * lus_sand_init; default 0): 0 simulate_sand (src/Simulate.cpp:156-325), the
 * default simulate_fun; 1 simulate_sand_credits (:327-510), installed by
 * init_simulation_extra_parameters(..., with_credits) (src/Lustrine.cpp:312):
 * attracted bit 2 marks a credit particle that falls at 1 until its first
 * contact; 2 simulate_sand_v3 (:513-667), wrapper index 1
 * (src/LustrineWrapper.cpp:666-677): dt clamped to [0.001, 0.016], no box
 * clamp, half-space walls, one contact pass.  Sets the collision, friction,
 * mu_s and mu_k constants to that function's locals.  A v3 prediction whose
 * cell index leaves the grid makes lus_sand_step fail with LUS_EINVAL. */
int lus_sand_set_variant(LusGpu* ctx, int variant);
/* State in the reference's host order (sorted slots of the last step).
 * perm (n_sand): the last counting-sort permutation, new slot -> old slot. */
int lus_sand_download(LusGpu* ctx, float* pos, float* pos_star, float* vel, int* attracted,
                      int* perm);
/* The last step's neighbor lists (find_neighbors_uniform_grid_v1,
 * src/neighbors/Neighbors.cpp:262-364), reference slot indices. */
int lus_sand_neighbors(LusGpu* ctx, long* off, int* idx, long cap, long* total);

/* --- scene construction --- */
/* init_chunk_from_grid (src/Lustrine.cpp:676-722) on the device: grid cells
 * X x Y x Z indexed x*Y*Z + y*Z + z (0 empty; with mask_out also cells
 * holding exactly 1 are skipped), each occupied cell subdivided into
 * subdivision^3 lattice points of spacing cell_size; point (x, y, z) of the
 * subdivided lattice, in x-major order, lands at
 * ((x * cell_size) + position) + cell_size / 2 per axis.  positions receives
 * num_occupied * subdivision^3 points (3 floats each; num_occupied is the
 * grid's num_occupied_grid_cells, which the chunk is sized by), the ones past
 * the generated count zero as in the reference's chunk; colors (nullable, 4
 * floats per point) the colour of each point's cell from cell_colors
 * (nullable, 4 floats per cell).  *generated receives the number of points;
 * more points than num_occupied * subdivision^3 is LUS_EINVAL (the reference
 * would write past its vector). */
int lus_chunk_from_grid(LusGpu* ctx, const int* cells, int X, int Y, int Z, const float* position, float cell_size,
                        int subdivision, int mask_out, int num_occupied, const float* cell_colors, float* positions,
                        float* colors, int* generated);
/* init_grid_from_magika_voxel (src/VoxelLoader.cpp:47-100) after the .vox file
 * is parsed: the model's sx*sy*sz colour indices (voxel (x, y, z) at
 * x + y*sx + z*sx*sy, 0 empty) become grid cells (x*sy*sz + y*sz + z) holding
 * the index, their colours (nullable, 4 floats per cell) palette[index] / 255
 * (palette: 256 rgba bytes), *occupied the number of set voxels. */
int lus_grid_from_voxels(LusGpu* ctx, const unsigned char* voxels, int sx, int sy, int sz,
                         const unsigned char* palette, int* cells, float* colors, int* occupied);

/* --- phase timing in the reference's profiling slots --- */
/* The reference keeps LUSTRINE_MAX_NUM_MEASUREMENTS = 7 counters holding the
 * last measurement of each phase and reads them with get_num_observation /
 * get_cycles / get_duration (src/profiling/Profiling.hpp:14-77, filled in
 * src/Lustrine.cpp:802-804 and src/Simulate.cpp:669-802).  With profiling on
 * for a context, each lus_sand_step call, and the last step of each
 * lus_pbf_step call, records hipEvents around its phases, waits for them, and
 * stores the GPU times in the same slots, process-wide:
 *   0  the step (the reference: simulate_fun, Bullet included)
 *   1  Bullet -- on the host, not measured here: 0
 *   2  the solver after Bullet (here: the whole GPU step)
 *   3  neighbour search (sand: sort + lists; PBF: predict, sort, tiles)
 *   4  constraint substeps (sand) / solver iterations with the list build,
 *      features and network (PBF)
 *   5  boundary collision -- inside the solver kernels here: 0
 *   6  velocity update (sand) / XSPH and the rerun decision (PBF)
 * Durations are seconds; cycles are duration x 2.7e9, the fixed rate the
 * reference's Unix build divides rdtsc cycles by, so cycles / 2.7e9 is the
 * duration as there.  Off by default (a profiled call waits for its
 * kernels). */
int       lus_profiling_enable(LusGpu* ctx, int on);
/* src/profiling/Profiling.hpp:75-77 with a lus_ prefix: slot count (7), and
 * the last measurement of slot `index` (0 outside 0..6). */
long      lus_get_num_observation(void);
long long lus_get_cycles(int index);
double    lus_get_duration(int index);

/* --- consumers of the grid state, served from the device cell tables --- */
/* query_cell_num_particles (src/Lustrine.cpp:864-914): the number of ids in
 * the cell lists of the last step's grid (uniform_gird_cells: each cell's
 * static solids, then its sand slots) over the cells from
 * get_cell_indices(min_pos) to get_cell_indices(max_pos) inclusive
 * (src/neighbors/Utils.hpp:13-21: clamp to [cs/2, domain - cs/2], divide,
 * truncate); include_solid == 0 counts only ids below the current sand count
 * (Simulation::ptr_sand_end).  The lists are the last step's, as in the
 * reference (a later sink eviction does not rebuild them); 0 before the first
 * step.  Only the count crosses to the host. */
int lus_sand_query_cells(LusGpu* ctx, const float* min_pos, const float* max_pos, int include_solid,
                         int* count);
/* Sink eviction (src/Lustrine.cpp:806-836): `cells` is the concatenation of
 * sink_cells[i] of every sink that fires this frame (state set and timer at
 * or above its frequency), in sink order.  Every sand slot listed by those
 * cells in the last step's grid is removed in list order as the reference
 * does it -- positions, positions_star, velocities and attracted of the last
 * sand slot move into the removed slot and the sand count shrinks by one --
 * so repeated cells and slots behave as there.  *n_sand (nullable) receives
 * the new sand count (the reference's ptr_sand_end); the caller keeps its
 * counters (num_sand_particles, num_remaining_sand_particles) and the sink
 * timers.  Only the listed cells' two cell starts cross to the host.  A cell
 * id outside the grid is LUS_EINVAL; before the first step nothing is
 * removed. */
int lus_sand_sink_evict(LusGpu* ctx, const int* cells, int ncells, int* n_sand);
/* Bullet coupling without a full download (src/BulletPhysics.cpp:602-646,
 * set_particles_box_colliders_positions over the sand range): the first
 * max_out sand slots, in slot order, whose current position x satisfies
 * |x - player| <= half_extent on every axis, where half_extent is what
 * particle_collide_with_player computes, (player_box_scale + 4 *
 * particle_radius) * 0.5 per axis.  slots[k] and pos[3k..3k+2] (each
 * nullable) receive the k-th such particle, *count how many were written
 * (at most max_out; the reference places num_particles_allocated = 100
 * colliders).  Only those records cross to the host. */
int lus_sand_player_box(LusGpu* ctx, const float* player, const float* half_extent, int max_out, int* slots,
                        float* pos, int* count);

/* ------------------------------------------------------------------ */
/* simulate_fluid (src/Simulate.cpp:27-115), the reference's fluid      */
/* Simulate_fun, with its sequential pressure sweep                    */
/* ------------------------------------------------------------------ */
/* The Simulation fields simulate_fluid reads (src/Simulation.hpp:141-171). */
typedef struct LusFluidParams {
    float particleRadius, kernelRadius, kernelFactor;
    float cubic_kernel_k, cubic_kernel_l;
    float mass, rest_density, relaxation_epsilon;
    float s_corr_dq, s_corr_k, s_corr_n;
    float gravity[3];
    float domain[3];   /* domainX, domainY, domainZ: resolve_collision uses (int) of each */
    LusGrid grid;      /* gridX/Y/Z and cell_size of find_neighbors_uniform_grid */
} LusFluidParams;

/* The state simulate_fluid works on, in the reference's slot layout: arrays of
 * `total` (Simulation::total_allocated) AoS float[3]; moving particles at
 * [sand_start, sand_end), solids at [solid_start, solid_end) (their
 * positions_star are what the lists test against; solids never move).
 * velocities are read at the moving slots only.  lambdas: the reference's
 * `lambdas` vector (n_lambdas entries; simulate_fluid reads lambdas[t] for the
 * t-th entry of a list, src/Simulate.cpp:97, so lists longer than n_lambdas
 * are reported by lus_fluid_step; it writes lambdas[i] for every moving slot,
 * so n_lambdas must reach sand_end). */
int lus_fluid_init(LusGpu* ctx, const LusFluidParams* p, int total, int sand_start, int sand_end,
                   int solid_start, int solid_end, const float* positions, const float* positions_star,
                   const float* velocities, int n_lambdas, const float* lambdas);
/* One simulate_fluid call (src/Simulate.cpp:27-115): dt clamped to [0.001,
 * 0.01], prediction, neighbour lists (Neighbors.cpp:366-450), lambdas, and the
 * pressure sweep in the reference's sequential order (particle i sees the
 * particles before it corrected), run level by level on the device.  Waits for
 * the device.  A predicted position whose cell index lies outside the grid
 * (the reference indexes outside its cell vector there) fails with LUS_EINVAL. */
int lus_fluid_step(LusGpu* ctx, float dt);
/* positions, positions_star (all `total` slots), velocities (moving slots),
 * lambdas (n_lambdas); each nullable. */
int lus_fluid_download(LusGpu* ctx, float* positions, float* positions_star, float* velocities, float* lambdas);
/* Of the last step: dependency levels of the sweep, list entries. */
int lus_fluid_stats(LusGpu* ctx, int* levels, long* list_entries);

#ifdef __cplusplus
}
#endif
#endif /* LUSTRINE_GPU_H */
