/*
 * lustrine_gpu_simulate.hpp -- drop-in Simulate_fun bodies over the C ABI.
 *
 * Header-only C++17.  The templates read and write a Lustrine::Simulation
 * (src/Simulation.hpp:135-251) by member name, so a maintainer instantiates
 * them inside the reference's own build, where Simulation is the real type:
 *
 *   void simulate_sand_gpu(Lustrine::Simulation* s, float dt) {
 *       Lustrine::Bullet::simulate_bullet(&s->bullet_physics_simulation, dt,
 *                                         s->ptr_sand_start, s->ptr_sand_end);
 *       lus_adapter::simulate_sand(s, dt);
 *   }
 *   sim.simulate_fun = simulate_sand_gpu;   // typedef at src/Simulation.hpp:22
 *
 * simulate_sand() replaces src/Simulate.cpp:156-325 (after its Bullet call):
 * it leaves positions, positions_star, velocities and attracted in the same
 * permuted slot order, with the same values, as the reference function.
 * simulate_pbf() is the position-based-fluids substep (no reference
 * counterpart with iterations and XSPH); particle order on the host is kept.
 *
 * query_cell_num_particles() and evict_sinks() serve the two host readers of
 * uniform_gird_cells (src/Lustrine.cpp:864-914 and :806-836) from the device
 * cell table, which the host lists are not filled from on this path.
 *
 * Device state persists per Simulation between calls.  It is uploaded again
 * when the sand range changes behind the adapter's back (sources move
 * ptr_sand_end, src/Lustrine.cpp:780; sinks run through evict_sinks() keep
 * it in step) or after mark_dirty().  Errors: the functions
 * return the LUS_E* code (0 = ok) and leave the host state untouched on
 * failure; the message is in lus_gpu_last_error().  Not thread-safe, like
 * the reference (one host thread per Simulation).
 */
#ifndef LUSTRINE_GPU_SIMULATE_HPP
#define LUSTRINE_GPU_SIMULATE_HPP

#include <cstring>
#include <unordered_map>
#include <vector>

#include "lustrine_gpu.h"

namespace lus_adapter {

struct State {
    LusGpu* ctx = nullptr;
    int kind = 0;  // 1 sand, 2 pbf
    int n_sand = -1;
    int ptr_solid_start = -1;
    bool dirty = true;
};

inline std::unordered_map<const void*, State>& states() {
    static std::unordered_map<const void*, State> m;
    return m;
}

/* Forces a full upload on the next call (host code edited the arrays). */
inline void mark_dirty(const void* sim) { states()[sim].dirty = true; }

/* Frees the device state of a Simulation (call before clean_simulation). */
inline void release(const void* sim) {
    auto it = states().find(sim);
    if (it == states().end()) return;
    lus_gpu_destroy(it->second.ctx);
    states().erase(it);
}

template <class V>
inline const float* fptr(const V* v) {
    static_assert(sizeof(V) == 3 * sizeof(float), "vec3 must be three packed floats");
    return reinterpret_cast<const float*>(v);
}
template <class V>
inline float* fptr(V* v) {
    static_assert(sizeof(V) == 3 * sizeof(float), "vec3 must be three packed floats");
    return reinterpret_cast<float*>(v);
}

inline int ensure_ctx(State& st, int device) {
    if (st.ctx) return LUS_OK;
    return lus_gpu_create(&st.ctx, device);
}

/* simulate_sand (src/Simulate.cpp:156-325) on the GPU. */
template <class Sim>
int simulate_sand(Sim* s, float dt, int device = 0) {
    State& st = states()[s];
    int rc = ensure_ctx(st, device);
    if (rc) return rc;
    const int n = s->ptr_sand_end - s->ptr_sand_start;
    if (s->ptr_sand_start != 0 || n <= 0) return LUS_EINVAL;
    if (st.kind != 1 || st.dirty || n != st.n_sand || s->ptr_solid_start != st.ptr_solid_start) {
        LusSandParams p{};
        p.particleRadius = s->particleRadius;
        p.particleDiameter = s->particleDiameter;
        p.kernelRadius = s->kernelRadius;
        p.mass = s->mass;
        p.domain[0] = s->domainX;
        p.domain[1] = s->domainY;
        p.domain[2] = s->domainZ;
        p.grid.cell_size = s->cell_size;
        p.grid.gridX = s->gridX;
        p.grid.gridY = s->gridY;
        p.grid.gridZ = s->gridZ;
        p.gravity[0] = s->gravity.x;
        p.gravity[1] = s->gravity.y;
        p.gravity[2] = s->gravity.z;
        /* the locals of src/Simulate.cpp:159-163 and its 4 substeps (:226) */
        p.collision_coeff = 0.8f;
        p.friction_coeff = 0.7f;
        p.mu_s = 0.95f;
        p.mu_k = 0.9f;
        p.substeps = 4;
        p.attract_radius = s->attract_radius;
        p.blow_radius = s->blow_radius;
        p.attract_coeff = s->attract_coeff;
        p.blow_coeff = s->blow_coeff;
        p.ptr_solid_start = s->ptr_solid_start - s->ptr_sand_start;
        const int ns = (s->ptr_solid_end > s->ptr_solid_start) ? s->ptr_solid_end - s->ptr_solid_start : 0;
        std::vector<int> solid_index((size_t)ns);
        for (int k = 0; k < ns; ++k) solid_index[k] = s->ptr_solid_start + k;
        const float* solid = ns ? fptr(s->positions_star + s->ptr_solid_start) : nullptr;
        rc = lus_sand_init(st.ctx, &p, n, fptr(s->positions), fptr(s->velocities), s->attracted, ns,
                           solid, ns ? solid_index.data() : nullptr);
        if (rc) return rc;
        st.kind = 1;
        st.n_sand = n;
        st.ptr_solid_start = s->ptr_solid_start;
        st.dirty = false;
    }
    const float player[3] = {s->bullet_physics_simulation.player_position.x,
                             s->bullet_physics_simulation.player_position.y,
                             s->bullet_physics_simulation.player_position.z};
    rc = lus_sand_step(st.ctx, dt, s->attract_flag ? 1 : 0, s->blow_flag ? 1 : 0, player);
    if (rc) return rc;
    rc = lus_sand_download(st.ctx, fptr(s->positions), fptr(s->positions_star), fptr(s->velocities),
                           s->attracted, nullptr);
    if (rc) return rc;
    s->time_step = dt;
    s->first_iteration = false;
    s->computed_static_particles = true;
    return LUS_OK;
}

/* Position-based-fluids substep over the sand range treated as fluid.
 * iterations: Jacobi solver iterations; exact: see LusPbfParams::exact. */
template <class Sim>
int simulate_pbf(Sim* s, float dt, int iterations = 4, int exact = 0, int device = 0) {
    State& st = states()[s];
    int rc = ensure_ctx(st, device);
    if (rc) return rc;
    const int n = s->ptr_sand_end - s->ptr_sand_start;
    if (s->ptr_sand_start != 0 || n <= 0) return LUS_EINVAL;
    if (st.kind != 2 || st.dirty || n != st.n_sand) {
        LusPbfParams p{};
        p.h = s->kernelRadius;
        p.f = s->kernelFactor;
        p.cell_size = s->cell_size;
        p.gridX = s->gridX;
        p.gridY = s->gridY;
        p.gridZ = s->gridZ;
        const float D[3] = {s->domainX, s->domainY, s->domainZ};
        for (int a = 0; a < 3; ++a) {
            p.lo[a] = s->particleRadius;
            p.hi[a] = D[a] - s->particleRadius;
        }
        p.gravity[0] = s->gravity.x;
        p.gravity[1] = s->gravity.y;
        p.gravity[2] = s->gravity.z;
        p.mass = s->mass;
        p.rest_density = s->rest_density;
        p.epsilon = s->relaxation_epsilon;
        p.scorr_k = s->s_corr_k;
        p.scorr_dq = s->s_corr_dq;
        p.scorr_n = (int)s->s_corr_n;
        p.c_xsph = s->c_xsph;
        p.iterations = iterations;
        p.exact = exact;
        rc = lus_pbf_init(st.ctx, &p, n, fptr(s->positions), fptr(s->velocities));
        if (rc) return rc;
        st.kind = 2;
        st.n_sand = n;
        st.dirty = false;
    }
    rc = lus_pbf_step(st.ctx, dt, 1);
    if (rc) return rc;
    rc = lus_pbf_download(st.ctx, 1, fptr(s->positions), fptr(s->velocities), nullptr);
    if (rc) return rc;
    std::memcpy(fptr(s->positions_star), fptr(s->positions), sizeof(float) * 3 * (size_t)n);
    s->time_step = dt;
    return LUS_OK;
}

/* query_cell_num_particles (src/Lustrine.cpp:864-914) for a Simulation
 * stepped by simulate_sand() above, answered from the device cell table of
 * the last step (the host uniform_gird_cells is not filled on this path).
 * *count receives the reference's return value. */
template <class Sim, class V>
int query_cell_num_particles(Sim* s, const V& min_pos, const V& max_pos, bool include_solid, int* count) {
    auto it = states().find(s);
    if (it == states().end() || it->second.kind != 1) return LUS_ESTATE;
    const float lo[3] = {min_pos.x, min_pos.y, min_pos.z};
    const float hi[3] = {max_pos.x, max_pos.y, max_pos.z};
    return lus_sand_query_cells(it->second.ctx, lo, hi, include_solid ? 1 : 0, count);
}

/* The sink loop of Lustrine::simulate (src/Lustrine.cpp:806-836) after
 * simulate_sand(): the cells of every firing sink (state set, timer at or
 * above its frequency; timers are advanced by the caller as there) are
 * evicted on the device in the reference's order, then the sand range of the
 * host arrays and the counters ptr_sand_end, num_sand_particles and
 * num_remaining_sand_particles are brought up to date.  Slots past the new
 * sand end are not written on the host. */
template <class Sim>
int evict_sinks(Sim* s) {
    auto it = states().find(s);
    if (it == states().end() || it->second.kind != 1) return LUS_ESTATE;
    State& st = it->second;
    std::vector<int> cells;
    for (int i = 0; i < s->sink->num_sinks; i++)
        if (s->sink->state[i] && s->sink->timers[i] >= s->sink->frequencies[i])
            cells.insert(cells.end(), s->sink->sink_cells[i].begin(), s->sink->sink_cells[i].end());
    if (cells.empty()) return LUS_OK;
    int n = st.n_sand;
    int rc = lus_sand_sink_evict(st.ctx, cells.data(), (int)cells.size(), &n);
    if (rc) return rc;
    const int removed = st.n_sand - n;
    if (removed == 0) return LUS_OK;
    if (n > 0) {
        rc = lus_sand_download(st.ctx, fptr(s->positions), fptr(s->positions_star), fptr(s->velocities),
                               s->attracted, nullptr);
        if (rc) return rc;
    }
    s->ptr_sand_end -= removed;
    s->num_sand_particles -= removed;
    s->num_remaining_sand_particles += removed;
    st.n_sand = n;
    return LUS_OK;
}

/* The particle side of set_particles_box_colliders_positions
 * (src/BulletPhysics.cpp:615-646) for a Simulation stepped by
 * simulate_sand(): the positions of the first num_particles_allocated sand
 * particles, in slot order, inside the player's box
 * (particle_collide_with_player, :602-606, around
 * bullet_physics_simulation.player_position, which the caller sets first as
 * the reference does) go to out[0..*count); the caller places that many
 * collider bodies and disables the rest.  The positions are the device's
 * current ones, i.e. the host arrays simulate_sand() left.  LUS_ESTATE before
 * the first simulate_sand() of this Simulation (use the host loop then). */
template <class Sim, class V>
int player_box_particles(Sim* s, V* out, int* count) {
    auto it = states().find(s);
    if (it == states().end() || it->second.kind != 1) return LUS_ESTATE;
    const auto& b = s->bullet_physics_simulation;
    const float r4 = b.particle_radius * 4.0f;
    const float half[3] = {(b.player_box_scale.x + r4) * 0.5f, (b.player_box_scale.y + r4) * 0.5f,
                           (b.player_box_scale.z + r4) * 0.5f};
    const float player[3] = {b.player_position.x, b.player_position.y, b.player_position.z};
    return lus_sand_player_box(it->second.ctx, player, half, b.num_particles_allocated, nullptr, fptr(out), count);
}

}  // namespace lus_adapter

#endif /* LUSTRINE_GPU_SIMULATE_HPP */
