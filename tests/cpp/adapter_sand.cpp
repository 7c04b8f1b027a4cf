// Drives the lus_adapter Simulate_fun bodies through a stand-in with the
// member names and layout of Lustrine::Simulation (src/Simulation.hpp:135-251)
// that the adapter reads.  Input/output: raw little-endian files.
//   adapter_sand.exe sand     <in> <out> <steps> <X> <Y> <Z> <kernel_radius_scale>
//   adapter_sand.exe credits  ... (simulate_sand_credits; attracted = 2 on every third slot)
//   adapter_sand.exe v3       ... (simulate_sand_v3)
//   adapter_sand.exe sandlazy ... (simulate_sand with the host arrays synchronised on demand)
//   adapter_sand.exe fluid    ... (simulate_fluid; out also float32[total] lambdas)
//   adapter_sand.exe pbf      <in> <out> <steps> <X> <Y> <Z> <kernel_radius_scale>
//   adapter_sand.exe bench    <in> <out> <steps> ...: times simulate_sand calls, every-step
//                             sync then on-demand sync, and writes "ms_every ms_lazy ms_sync"
//   adapter_sand.exe sandsink <in> <out> <steps> <X> <Y> <Z> <kernel_radius_scale>
//   adapter_sand.exe sandbox  <in> <out> <steps> <X> <Y> <Z> <kernel_radius_scale>
// in:  int32 n_sand, int32 n_solid, int32 total, float32[3*n_sand] sand,
//      float32[3*n_solid] solids (written downward from slot total-1);
//      sandsink: then int32 ncells, int32[ncells] cells of one sink that fires
//      after every step (the sink loop of Lustrine::simulate)
// out: float32[3*n] positions, float32[3*n] velocities, int32[n] attracted;
//      sandsink: first int32 n (the final ptr_sand_end), the whole-domain
//      query_cell_num_particles without and with solids, and
//      num_remaining_sand_particles;
//      sandbox: then float32 player_position[3], player_box_scale[3]; out
//      first int32 count and float32[3*count] collider positions
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "lustrine_gpu_simulate.hpp"

namespace stand_in {
struct vec3 {
    float x, y, z;
};
struct BulletSim {  // src/BulletPhysicsSimulation.hpp: the members the adapter reads
    vec3 player_position{0.0f, 0.0f, 0.0f};
    int num_particles_allocated = 100;
    vec3 player_box_scale{1.0f, 2.0f, 1.0f};
    float particle_radius = 0.1f;
};
struct ParticleSink {  // src/Simulation.hpp:94-103
    std::vector<bool> state;
    std::vector<float> timers;
    std::vector<float> frequencies;
    std::vector<int> despawned;
    std::vector<std::vector<int>> sink_cells;
    int num_sinks = 0;
    std::vector<int> temp_removal;
};
struct Simulation {  // members read by the adapter, with their reference types
    BulletSim bullet_physics_simulation;
    float particleRadius, particleDiameter, kernelRadius, kernelFactor = 0.5f;
    float domainX, domainY, domainZ;
    float rest_density = 24.0f, mass = 5.0f;
    vec3 gravity{0.0f, -10.0f, 0.0f};
    float time_step = 0.01f, relaxation_epsilon = 10.0f;
    float s_corr_dq = 0.5f, s_corr_k = 1.0f, s_corr_n = 4, c_xsph = 0.1f;
    vec3* velocities = nullptr;
    std::vector<float> lambdas;
    float cubic_kernel_k = 0.0f, cubic_kernel_l = 0.0f;
    size_t total_allocated = 0;
    int gridX, gridY, gridZ;
    float cell_size;
    bool computed_static_particles = false;
    vec3* positions = nullptr;
    vec3* positions_star = nullptr;
    int ptr_sand_start = -1, ptr_sand_end = -1, ptr_solid_start = -1, ptr_solid_end = -1;
    int num_sand_particles = 0;
    int num_remaining_sand_particles = 0;
    ParticleSink* sink = nullptr;
    bool first_iteration = true;
    bool attract_flag = false, blow_flag = false;
    int* attracted = nullptr;
    float attract_radius = 1.5f, blow_radius = 2.0f, attract_coeff = 1000.0f, blow_coeff = 500.0f;
};
}  // namespace stand_in

int main(int argc, char** argv) {
    if (argc != 9) {
        fprintf(stderr, "usage: %s sand|pbf in out steps X Y Z scale\n", argv[0]);
        return 2;
    }
    const std::string mode = argv[1];
    FILE* f = fopen(argv[2], "rb");
    if (!f) return 2;
    int hdr[3];
    if (fread(hdr, sizeof(int), 3, f) != 3) return 2;
    const int n = hdr[0], ns = hdr[1], total = hdr[2];
    std::vector<stand_in::vec3> pos((size_t)total), ps((size_t)total), vel((size_t)total);
    std::vector<int> att((size_t)total, 0);
    std::vector<float> buf(3 * (size_t)(n + ns));
    if (fread(buf.data(), sizeof(float), buf.size(), f) != buf.size()) return 2;
    stand_in::ParticleSink sink;
    float box_in[6] = {0, 0, 0, 0, 0, 0};
    if (mode == "sandbox" && fread(box_in, sizeof(float), 6, f) != 6) return 2;
    if (mode == "sandsink") {
        int nc = 0;
        if (fread(&nc, sizeof(int), 1, f) != 1) return 2;
        std::vector<int> cells((size_t)nc);
        if (fread(cells.data(), sizeof(int), (size_t)nc, f) != (size_t)nc) return 2;
        sink.num_sinks = 1;
        sink.state = {true};
        sink.timers = {0.0f};
        sink.frequencies = {0.0f};  // fires on every frame
        sink.despawned = {0};
        sink.sink_cells = {cells};
    }
    fclose(f);
    for (int i = 0; i < n; ++i) pos[i] = {buf[3 * i], buf[3 * i + 1], buf[3 * i + 2]};
    for (int k = 0; k < ns; ++k) {
        const float* q = &buf[3 * (size_t)(n + k)];
        pos[total - 1 - k] = {q[0], q[1], q[2]};  // solids written downward (src/Lustrine.cpp:491-504)
    }
    ps = pos;
    stand_in::Simulation s;
    const int X = atoi(argv[5]), Y = atoi(argv[6]), Z = atoi(argv[7]);
    const float scale = (float)atof(argv[8]);
    s.particleRadius = 0.5f;
    s.particleDiameter = 1.0f;
    s.domainX = (float)X;
    s.domainY = (float)Y;
    s.domainZ = (float)Z;
    s.kernelRadius = scale * s.particleRadius;  // src/Lustrine.cpp:545-546
    s.cell_size = 1.0f * s.kernelRadius;
    s.gridX = (int)(s.domainX / s.cell_size) + 1;
    s.gridY = (int)(s.domainY / s.cell_size) + 1;
    s.gridZ = (int)(s.domainZ / s.cell_size) + 1;
    {   // src/Lustrine.cpp:253-259: k and l in double from h^3 as a float
        const double pi = 3.14159265358979323846;
        const float h3 = (float)std::pow((double)s.kernelRadius, 3.0);
        s.cubic_kernel_k = (float)(8.0f / (pi * (double)h3));
        s.cubic_kernel_l = (float)(48.0f / (pi * (double)h3));
    }
    s.total_allocated = (size_t)total;
    s.lambdas.assign((size_t)total, 0.0f);
    if (mode == "credits")
        for (int i = 0; i < n; i += 3) att[i] = 2;
    s.positions = pos.data();
    s.positions_star = ps.data();
    s.velocities = vel.data();
    s.attracted = att.data();
    s.ptr_sand_start = 0;
    s.ptr_sand_end = n;
    s.num_sand_particles = n;
    s.ptr_solid_start = ns ? total - ns : total - 1;
    s.ptr_solid_end = ns ? total : total - 1;
    s.sink = &sink;
    const int steps = atoi(argv[4]);
    if (mode == "bench") {
        // one warm call, then `steps` timed calls with each sync mode
        auto now = [] { return std::chrono::steady_clock::now(); };
        auto ms = [](auto a, auto b) { return std::chrono::duration<double, std::milli>(b - a).count(); };
        if (lus_adapter::simulate_sand(&s, 0.016f)) return 1;
        auto t0 = now();
        for (int t = 0; t < steps; ++t)
            if (lus_adapter::simulate_sand(&s, 0.016f)) return 1;
        auto t1 = now();
        lus_adapter::set_sync(&s, lus_adapter::kSyncOnDemand);
        auto t2 = now();
        for (int t = 0; t < steps; ++t)
            if (lus_adapter::simulate_sand(&s, 0.016f)) return 1;
        lus_gpu_sync(lus_adapter::states()[&s].ctx);
        auto t3 = now();
        if (lus_adapter::sync_host(&s)) return 1;
        auto t4 = now();
        lus_adapter::release(&s);
        FILE* o = fopen(argv[3], "w");
        if (!o) return 2;
        fprintf(o, "%.4f %.4f %.4f\n", ms(t0, t1) / steps, ms(t2, t3) / steps, ms(t3, t4));
        fclose(o);
        return 0;
    }
    if (mode == "sandlazy") lus_adapter::set_sync(&s, lus_adapter::kSyncOnDemand);
    for (int t = 0; t < steps; ++t) {
        int rc;
        if (mode == "pbf") rc = lus_adapter::simulate_pbf(&s, 1.0f / 60.0f, 4, 1);
        else if (mode == "credits") rc = lus_adapter::simulate_sand_credits(&s, 0.016f);
        else if (mode == "v3") rc = lus_adapter::simulate_sand_v3(&s, 0.016f);
        else if (mode == "fluid") rc = lus_adapter::simulate_fluid(&s, 1.0f / 60.0f);
        else rc = lus_adapter::simulate_sand(&s, 0.016f);
        if (!rc && mode == "sandsink") rc = lus_adapter::evict_sinks(&s);
        if (rc) {
            char msg[512];
            lus_gpu_last_error(msg, sizeof msg);
            fprintf(stderr, "step %d failed: %s\n", t, msg);
            return 1;
        }
    }
    if (mode == "sandlazy") {
        // positions only first (what simulation_bind_positions_copy reads),
        // then the rest
        if (lus_adapter::sync_positions(&s)) return 1;
        if (lus_adapter::states()[&s].pos_stale || !lus_adapter::states()[&s].rest_stale) return 3;
        if (lus_adapter::sync_host(&s)) return 1;
    }
    int q[2] = {0, 0};
    if (mode == "sandsink") {
        const stand_in::vec3 lo{0.0f, 0.0f, 0.0f}, hi{s.domainX, s.domainY, s.domainZ};
        if (lus_adapter::query_cell_num_particles(&s, lo, hi, false, &q[0]) ||
            lus_adapter::query_cell_num_particles(&s, lo, hi, true, &q[1]))
            return 1;
    }
    std::vector<stand_in::vec3> near((size_t)s.bullet_physics_simulation.num_particles_allocated);
    int near_count = 0;
    if (mode == "sandbox") {
        s.bullet_physics_simulation.player_position = {box_in[0], box_in[1], box_in[2]};
        s.bullet_physics_simulation.player_box_scale = {box_in[3], box_in[4], box_in[5]};
        if (lus_adapter::player_box_particles(&s, near.data(), &near_count)) return 1;
    }
    lus_adapter::release(&s);
    FILE* o = fopen(argv[3], "wb");
    if (!o) return 2;
    if (mode == "sandbox") {
        fwrite(&near_count, sizeof(int), 1, o);
        fwrite(near.data(), sizeof(float), 3 * (size_t)near_count, o);
    }
    if (mode == "sandsink") {
        const int hdr_out[4] = {s.ptr_sand_end, q[0], q[1], s.num_remaining_sand_particles};
        fwrite(hdr_out, sizeof(int), 4, o);
    }
    const int nw = mode == "sandsink" ? s.ptr_sand_end : n;
    fwrite(pos.data(), sizeof(float), 3 * (size_t)nw, o);
    fwrite(vel.data(), sizeof(float), 3 * (size_t)nw, o);
    fwrite(att.data(), sizeof(int), (size_t)nw, o);
    if (mode == "fluid") {
        fwrite(ps.data(), sizeof(float), 3 * (size_t)total, o);
        fwrite(s.lambdas.data(), sizeof(float), (size_t)total, o);
    }
    fclose(o);
    return 0;
}
