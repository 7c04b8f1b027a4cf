/*
 * lus_oracle.h -- CPU oracle for the Lustrine particle substep.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (lustrine_amd/, include/)
 * links, loads or calls this code.  It is used by tests/ as the checker, by
 * __graft_entry__.smoke() as the checker, and by bench.py's cpu_baseline leg.
 *
 * Two families of functions:
 *   ref_*  : restatement of the reference's own arithmetic and loop order
 *            (GrapixLeGrand/Lustrine, files cited per function), built with
 *            -O2 -fno-fast-math -ffp-contract=off so every float op is a
 *            single IEEE-754 binary32 (or binary64 where the reference promotes).
 *   pbf_*  : the position-based-fluids substep with poly6/spiky kernels,
 *            K Jacobi iterations and XSPH (the physics the GPU path computes),
 *            in an FP32 variant (same operation order as the GPU kernels) and
 *            an FP64 twin used to state the FP32 deviation.
 *
 * Vectors are AoS float[3] (the layout of glm::vec3 in the reference).
 */
#ifndef LUS_ORACLE_H
#define LUS_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Fields of Lustrine::Simulation used on the path (src/Simulation.hpp:143-170). */
typedef struct OrcSim {
    float particleRadius;
    float particleDiameter;
    float kernelRadius;
    float kernelFactor;
    float cubic_kernel_k;
    float cubic_kernel_l;
    float domainX, domainY, domainZ;
    float rest_density;
    float mass;
    float relaxation_epsilon;
    float s_corr_dq, s_corr_k, s_corr_n;
    float c_xsph;
    float gravity[3];
    float cell_size;
    int gridX, gridY, gridZ, num_grid_cells;
} OrcSim;

/* src/Lustrine.cpp:383-386, 543-558 and the member defaults of src/Simulation.hpp:147-170. */
void orc_init_constants(OrcSim* s, int X, int Y, int Z, float particleRadius,
                        float particleDiameter, float kernel_radius_scale);

/* ---- SPH kernels, src/Kernels.cpp ---- */
float orc_cubic_kernel(const OrcSim* s, float r);                        /* :6-19  */
void  orc_cubic_kernel_grad(const OrcSim* s, const float r[3], float out[3]); /* :26-41 */
float orc_poly6_kernel(const OrcSim* s, float r);                        /* :43-51 */
float orc_poly6_kernel_vec(const OrcSim* s, const float r[3]);           /* :53-55 */
void  orc_spiky_kernel(const OrcSim* s, const float r[3], float out[3]); /* :57-67 */

/* ---- grid, src/neighbors ---- */
void orc_cell_ids(const OrcSim* s, const float* pos, int n, int* out);   /* Utils.hpp:24-33 */
void orc_counting_sort(int* counts, const int* keys, size_t n, size_t num_cells,
                       int* sorted);                                     /* Sorting.cpp:11-33 */

/* Fluid-path neighbor lists, src/neighbors/Neighbors.cpp:366-450.
 * positions_star has `total` slots; sand = [sand_start, sand_end),
 * solid = [solid_start, solid_end).  Output CSR over sand slots:
 * off[k]..off[k+1] for slot sand_start+k.  Returns the number of entries, or
 * -(needed) if cap is too small (nothing valid written then). */
long orc_neighbors_fluid(const OrcSim* s, const float* positions_star,
                         int sand_start, int sand_end, int solid_start, int solid_end,
                         long* off, int* idx, long cap);

/* State of one reference simulate_sand stream (src/Simulate.cpp:156-325). */
typedef struct OrcSandState {
    size_t total_allocated;
    int ptr_sand_start, ptr_sand_end, ptr_solid_start, ptr_solid_end;
    int num_sand_particles;
    float* positions;       /* [3*total] */
    float* positions_star;  /* [3*total] */
    float* positions_tmp;   /* [3*total] */
    float* velocities;      /* [3*total] */
    int*   attracted;       /* [total]   */
    int first_iteration;
    int computed_static_particles;
    int attract_flag, blow_flag, prev_attract_flag;
    float player_position[3];
    float attract_radius, blow_radius, attract_coeff, blow_coeff;
    /* solid cell cache (uniform_grid_cells_static_saved), CSR over cells, owned */
    long* static_off;
    int*  static_idx;
    /* outputs of the last step: permutation (new slot -> old slot) of the
     * counting sort and the neighbor lists, CSR over sand slots, owned */
    int*  sorted_indices;
    long* nbr_off;
    int*  nbr_idx;
    long  nbr_cap;
    /* sand part of every cell's list in the last step's grid
     * (uniform_gird_cells after the static solids): cell c holds the sand
     * slots [cell_sand_off[c], cell_sand_off[c+1]); owned, NULL before the
     * first step */
    long* cell_sand_off;
} OrcSandState;

/* query_cell_num_particles (src/Lustrine.cpp:864-914) over the last step's
 * cell lists; 0 before the first step (the lists are empty then). */
int orc_query_cell_num_particles(const OrcSim* s, const OrcSandState* st, const float* min_pos,
                                 const float* max_pos, int include_solid);
/* Sink eviction (src/Lustrine.cpp:806-836) for the concatenated cell lists
 * `cells` of the firing sinks.  Returns the number of particles removed. */
int orc_sink_evict(const OrcSim* s, OrcSandState* st, const int* cells, int ncells);

/* One call of simulate_sand.  Returns 0 on success. */
int  orc_simulate_sand(const OrcSim* s, OrcSandState* st, float dt);
/* One call of the sand function `variant`: 0 simulate_sand (src/Simulate.cpp:
 * 156-325), 1 simulate_sand_credits (:327-510), 2 simulate_sand_v3 (:513-667). */
int  orc_simulate_sand_variant(const OrcSim* s, OrcSandState* st, float dt, int variant);
void orc_sand_state_free(OrcSandState* st);

/* One call of simulate_fluid (src/Simulate.cpp:27-115), sequential in-place
 * update as the reference.  lambdas has sand_end slots.  Returns 0 on success. */
int orc_simulate_fluid(OrcSim* s, float* positions, float* positions_star,
                       float* velocities, float* lambdas,
                       int sand_start, int sand_end, int solid_start, int solid_end,
                       float dt);

/* ---- position-based fluids substep (the path the GPU computes) ---- */
typedef struct OrcPbfParams {
    float h;            /* kernel radius */
    float f;            /* kernelFactor */
    float cell_size;
    int gridX, gridY, gridZ;
    float lo[3], hi[3]; /* box clamp bounds: [r, D - r] per axis */
    float gravity[3];
    float mass, rest_density, epsilon;
    float scorr_k, scorr_dq; int scorr_n;
    float c_xsph;
    int iterations;
} OrcPbfParams;

/* One substep over n fluid particles.  x, v: [3n] in/out, id: [n] in/out
 * (the carried original index).  State comes back in cell order.
 * max_c (optional): max over particles of the density constraint C_i from the
 * last lambda pass.  final_err (optional, 2 values): max(0, max_i C_i) and
 * mean_i max(0, C_i) at the final positions over the step's neighbour sets
 * (summed in the XSPH loop). */
int orc_pbf_step_f32(const OrcPbfParams* p, int n, float* x, float* v, int* id,
                     float dt, float* max_c, float* final_err);
int orc_pbf_step_f64(const OrcPbfParams* p, int n, double* x, double* v, int* id,
                     double dt, double* max_c, double* final_err);

/* Density error of a state: max_i max(0, rho_i/rho0 - 1) with the grid built
 * from x itself (no predict).  Used to measure the learned warm start. */
double orc_pbf_density_error_f64(const OrcPbfParams* p, int n, const double* x);

/* ---- mixed sand + fluid substep over static solid particles ---- */
typedef struct OrcMixParams {
    float diameter;                        /* contact distance d = 2r */
    float collision, friction, mu_s, mu_k; /* src/Simulate.cpp:159-163 roles */
} OrcMixParams;

/* One substep; phase[i]: 0 fluid, 1 sand, 2 solid (static).  x, v, id, phase
 * in/out, returned in cell order; p->iterations Jacobi iterations.
 * final_err (optional, 2 values): max and mean over fluid particles of
 * max(0, rho_i/rho0 - 1) at the final positions.  See mixed_core.inc. */
int orc_mixed_step_f32(const OrcPbfParams* p, const OrcMixParams* mp, int n, float* x, float* v,
                       int* id, int* phase, float dt, float* final_err);
int orc_mixed_step_f64(const OrcPbfParams* p, const OrcMixParams* mp, int n, double* x, double* v,
                       int* id, int* phase, double dt, double* final_err);

#ifdef __cplusplus
}
#endif
#endif
