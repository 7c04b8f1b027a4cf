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
     *    neighbour loops (results within a few ulp of mode 1).  The
     *    neighbour sets are built with the exact test in both modes. */
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
/* Enqueues `steps` substeps of length dt on the context stream. */
int lus_pbf_step(LusGpu* ctx, float dt, int steps);
/* on = 1: lus_pbf_step replays a hipGraph holding two captured substeps
 * (captured on first use and again when dt changes); on = 0: plain launches. */
int lus_pbf_set_graph(LusGpu* ctx, int on);
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
 * (array of LUS_PBF_NPHASES) the `steps` instead record an event after every
 * kernel: phase_ms accumulates per-phase device time, phase_launches the
 * launch count per phase, and *ms their sum. */
#define LUS_PBF_NPHASES 8
/* phases: 0 predict+cell key, 1 cell scan, 2 scatter, 3 stable permute,
 *         4 lambda pass (iterations 2..K), 5 delta-p pass, 6 XSPH + commit,
 *         7 first lambda pass of the step (builds the neighbour sets) */
int lus_pbf_bench(LusGpu* ctx, float dt, int warmup, int steps, float* ms,
                  float* phase_ms, int* phase_launches);
/* Average device time per launch (microseconds) of the neighbour-pass kernels
 * on the current state: us[0] lambda pass, us[1] delta-p pass, us[2] XSPH
 * pass, us[3] the first lambda pass of a step (which also builds the
 * neighbour sets), each launched `reps` times back to back between two
 * hipEvents on the context stream.  Outputs go to scratch buffers; the scene
 * state is left unchanged.  Requires at least one step since init. */
int lus_pbf_time_kernels(LusGpu* ctx, int reps, float* us);

/* --- sand scene: simulate_sand (src/Simulate.cpp:156-325) --- */
/* n_sand moving particles (pos, vel, attracted may be NULL for zero) and
 * n_solid static particles with their reference slots solid_index[k]. */
int lus_sand_init(LusGpu* ctx, const LusSandParams* p, int n_sand, const float* pos,
                  const float* vel, const int* attracted, int n_solid,
                  const float* solid_pos, const int* solid_index);
/* One simulate_sand call.  player = Bullet player position (3 floats, may be
 * NULL when both flags are 0). */
int lus_sand_step(LusGpu* ctx, float dt, int attract_flag, int blow_flag, const float* player);
/* State in the reference's host order (sorted slots of the last step).
 * perm (n_sand): the last counting-sort permutation, new slot -> old slot. */
int lus_sand_download(LusGpu* ctx, float* pos, float* pos_star, float* vel, int* attracted,
                      int* perm);
/* The last step's neighbor lists (find_neighbors_uniform_grid_v1,
 * src/neighbors/Neighbors.cpp:262-364), reference slot indices. */
int lus_sand_neighbors(LusGpu* ctx, long* off, int* idx, long cap, long* total);

#ifdef __cplusplus
}
#endif
#endif /* LUSTRINE_GPU_H */
