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
 *       if (lus_adapter::simulate_sand(s, dt) != LUS_OK) {   // a HIP failure: logged
 *           lus_adapter::release(s);
 *           Lustrine::simulate_sand(s, dt);                   // the reference, this step
 *       }
 *   }
 *   sim.simulate_fun = simulate_sand_gpu;   // typedef at src/Simulation.hpp:22
 *
 * simulate_sand() replaces src/Simulate.cpp:156-325 (after its Bullet call),
 * simulate_sand_credits() :327-510 and simulate_sand_v3() :513-667: they leave
 * positions, positions_star, velocities and attracted in the same permuted
 * slot order, with the same values, as the reference functions.
 * simulate_fluid() replaces :27-115 (same values up to s_coor's pow, see
 * lus_fluid_step).  simulate_pbf() is the position-based-fluids substep (no
 * reference counterpart with iterations and XSPH); particle order on the host
 * is kept.
 *
 * Host arrays: by default every call leaves the results in the Simulation's
 * arrays, as the reference does.  set_sync(sim, kSyncOnDemand) keeps them on
 * the device instead; sync_positions() (what simulation_bind_positions_copy
 * reads) or sync_host() (every array) bring them back, and must be called
 * before host code reads or edits the arrays.
 *
 * query_cell_num_particles() and evict_sinks() serve the two host readers of
 * uniform_gird_cells (src/Lustrine.cpp:864-914 and :806-836) from the device
 * cell table, which the host lists are not filled from on this path.
 *
 * Device state persists per Simulation between calls.  It is uploaded again
 * when the sand range changes behind the adapter's back (sources move
 * ptr_sand_end, src/Lustrine.cpp:780; sinks run through evict_sinks() keep
 * it in step) or after mark_dirty().  Errors: the functions return the LUS_E*
 * code (0 = ok) and leave the host arrays as the last synchronised state; the
 * message is in lus_gpu_last_error().  Not thread-safe, like the reference
 * (one host thread per Simulation).
 */
#ifndef LUSTRINE_GPU_SIMULATE_HPP
#define LUSTRINE_GPU_SIMULATE_HPP

#include <cstring>
#include <unordered_map>
#include <vector>

#include "lustrine_gpu.h"

namespace lus_adapter {

enum Sync { kSyncEveryStep = 0, kSyncOnDemand = 1 };

struct State {
    LusGpu* ctx = nullptr;
    int kind = 0;     // 1 sand, 2 pbf, 3 simulate_fluid
    int variant = 0;  // sand: 0 simulate_sand, 1 simulate_sand_credits, 2 simulate_sand_v3
    int n_sand = -1;
    int ptr_solid_start = -1;
    bool dirty = true;
    int sync = kSyncEveryStep;
    bool pos_stale = false;   // the device holds newer positions than the host
    bool rest_stale = false;  // ... newer velocities / attracted / lambdas
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

/* Brings the host arrays up to the device state (all of them, or positions
 * and positions_star only). */
template <class Sim>
int sync_host(Sim* s, bool positions_only = false) {
    auto it = states().find(s);
    if (it == states().end()) return LUS_OK;
    State& st = it->second;
    if (!st.pos_stale && (positions_only || !st.rest_stale)) return LUS_OK;
    int rc = LUS_OK;
    const bool rest = !positions_only && st.rest_stale;
    if (st.kind == 1) {
        rc = lus_sand_download(st.ctx, fptr(s->positions), fptr(s->positions_star),
                               rest ? fptr(s->velocities) : nullptr, rest ? s->attracted : nullptr, nullptr);
    } else if (st.kind == 2) {
        rc = lus_pbf_download(st.ctx, 1, fptr(s->positions), rest ? fptr(s->velocities) : nullptr, nullptr);
        if (rc == LUS_OK)
            std::memcpy(fptr(s->positions_star), fptr(s->positions), sizeof(float) * 3 * (size_t)st.n_sand);
    } else if (st.kind == 3) {
        rc = lus_fluid_download(st.ctx, fptr(s->positions), fptr(s->positions_star),
                                rest ? fptr(s->velocities) : nullptr, rest ? s->lambdas.data() : nullptr);
    }
    if (rc) return rc;
    st.pos_stale = false;
    if (rest) st.rest_stale = false;
    return LUS_OK;
}

/* What simulation_bind_positions_copy (src/LustrineWrapper.cpp:377-379) reads. */
template <class Sim>
int sync_positions(Sim* s) { return sync_host(s, true); }

/* kSyncEveryStep (default): each call leaves its results in the host arrays.
 * kSyncOnDemand: they stay on the device until sync_positions / sync_host. */
template <class Sim>
int set_sync(Sim* s, int mode) {
    int rc = LUS_OK;
    if (mode == kSyncEveryStep) rc = sync_host(s);
    states()[s].sync = mode;
    return rc;
}

inline void after_step(State& st) {
    if (st.sync == kSyncEveryStep) return;
    st.pos_stale = true;
    st.rest_stale = true;
}

/* The sand Simulate_funs on the GPU: variant 0 simulate_sand
 * (src/Simulate.cpp:156-325), 1 simulate_sand_credits (:327-510), 2
 * simulate_sand_v3 (:513-667). */
template <class Sim>
int simulate_sand_variant(Sim* s, float dt, int variant, int device = 0) {
    State& st = states()[s];
    int rc = ensure_ctx(st, device);
    if (rc) return rc;
    const int n = s->ptr_sand_end - s->ptr_sand_start;
    if (s->ptr_sand_start != 0 || n <= 0) return LUS_EINVAL;
    if (st.kind != 1 || st.dirty || n != st.n_sand || s->ptr_solid_start != st.ptr_solid_start) {
        // device results not yet on the host keep their slots below the old end
        if (st.kind && (st.pos_stale || st.rest_stale) && n >= st.n_sand) {
            rc = sync_host(s);
            if (rc) return rc;
        }
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
        /* the locals of src/Simulate.cpp:159-163 and its 4 substeps (:226);
         * lus_sand_set_variant below sets those of the other functions */
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
        st.variant = 0;
        st.n_sand = n;
        st.ptr_solid_start = s->ptr_solid_start;
        st.dirty = false;
        st.pos_stale = st.rest_stale = false;
    }
    if (st.variant != variant) {
        rc = lus_sand_set_variant(st.ctx, variant);
        if (rc) return rc;
        st.variant = variant;
    }
    const float player[3] = {s->bullet_physics_simulation.player_position.x,
                             s->bullet_physics_simulation.player_position.y,
                             s->bullet_physics_simulation.player_position.z};
    rc = lus_sand_step(st.ctx, dt, s->attract_flag ? 1 : 0, s->blow_flag ? 1 : 0, player);
    if (rc) return rc;
    if (st.sync == kSyncEveryStep) {
        rc = lus_sand_download(st.ctx, fptr(s->positions), fptr(s->positions_star), fptr(s->velocities),
                               s->attracted, nullptr);
        if (rc) return rc;
    }
    after_step(st);
    if (variant == 2) {  // src/Simulate.cpp:522-524
        const float t = dt < 0.001f ? 0.001f : dt;
        s->time_step = 0.016f < t ? 0.016f : t;
    } else {
        s->time_step = dt;
        s->first_iteration = false;
    }
    s->computed_static_particles = true;
    return LUS_OK;
}

template <class Sim>
int simulate_sand(Sim* s, float dt, int device = 0) { return simulate_sand_variant(s, dt, 0, device); }
template <class Sim>
int simulate_sand_credits(Sim* s, float dt, int device = 0) { return simulate_sand_variant(s, dt, 1, device); }
template <class Sim>
int simulate_sand_v3(Sim* s, float dt, int device = 0) { return simulate_sand_variant(s, dt, 2, device); }

/* simulate_fluid (src/Simulate.cpp:27-115) on the GPU: every moving slot's
 * positions, positions_star, velocities and the lambdas vector as the
 * reference leaves them (slot order unchanged, as the reference keeps it). */
template <class Sim>
int simulate_fluid(Sim* s, float dt, int device = 0) {
    State& st = states()[s];
    int rc = ensure_ctx(st, device);
    if (rc) return rc;
    const int n = s->ptr_sand_end - s->ptr_sand_start;
    if (n < 0) return LUS_EINVAL;
    if (st.kind != 3 || st.dirty || n != st.n_sand || s->ptr_solid_start != st.ptr_solid_start) {
        if (st.kind && (st.pos_stale || st.rest_stale) && n >= st.n_sand) {
            rc = sync_host(s);
            if (rc) return rc;
        }
        LusFluidParams p{};
        p.particleRadius = s->particleRadius;
        p.kernelRadius = s->kernelRadius;
        p.kernelFactor = s->kernelFactor;
        p.cubic_kernel_k = s->cubic_kernel_k;
        p.cubic_kernel_l = s->cubic_kernel_l;
        p.mass = s->mass;
        p.rest_density = s->rest_density;
        p.relaxation_epsilon = s->relaxation_epsilon;
        p.s_corr_dq = s->s_corr_dq;
        p.s_corr_k = s->s_corr_k;
        p.s_corr_n = s->s_corr_n;
        p.gravity[0] = s->gravity.x;
        p.gravity[1] = s->gravity.y;
        p.gravity[2] = s->gravity.z;
        p.domain[0] = s->domainX;
        p.domain[1] = s->domainY;
        p.domain[2] = s->domainZ;
        p.grid.cell_size = s->cell_size;
        p.grid.gridX = s->gridX;
        p.grid.gridY = s->gridY;
        p.grid.gridZ = s->gridZ;
        rc = lus_fluid_init(st.ctx, &p, (int)s->total_allocated, s->ptr_sand_start, s->ptr_sand_end,
                            s->ptr_solid_start, s->ptr_solid_end, fptr(s->positions), fptr(s->positions_star),
                            fptr(s->velocities), (int)s->lambdas.size(), s->lambdas.data());
        if (rc) return rc;
        st.kind = 3;
        st.n_sand = n;
        st.ptr_solid_start = s->ptr_solid_start;
        st.dirty = false;
        st.pos_stale = st.rest_stale = false;
    }
    rc = lus_fluid_step(st.ctx, dt);
    if (rc) return rc;
    if (st.sync == kSyncEveryStep) {
        rc = lus_fluid_download(st.ctx, fptr(s->positions), fptr(s->positions_star), fptr(s->velocities),
                                s->lambdas.data());
        if (rc) return rc;
    }
    after_step(st);
    const float t = dt < 0.001f ? 0.001f : dt;  // src/Simulate.cpp:31-32
    s->time_step = 0.01f < t ? 0.01f : t;
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
        if (st.kind && (st.pos_stale || st.rest_stale) && n >= st.n_sand) {
            rc = sync_host(s);
            if (rc) return rc;
        }
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
        st.pos_stale = st.rest_stale = false;
    }
    rc = lus_pbf_step(st.ctx, dt, 1);
    if (rc) return rc;
    if (st.sync == kSyncEveryStep) {
        rc = lus_pbf_download(st.ctx, 1, fptr(s->positions), fptr(s->velocities), nullptr);
        if (rc) return rc;
        std::memcpy(fptr(s->positions_star), fptr(s->positions), sizeof(float) * 3 * (size_t)n);
    }
    after_step(st);
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
    if (n > 0 && st.sync == kSyncEveryStep) {
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
