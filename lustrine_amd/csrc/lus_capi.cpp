// lus_capi.cpp -- host side of the C ABI declared in include/lustrine_gpu.h:
// argument checks, device buffers, the per-step launch sequences and timing.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <limits>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <random>
#include <unordered_map>
#include <vector>

#include "lus_device.hpp"
#include "lus_internal.hpp"
#include "lustrine_gpu.h"

namespace {

int g_err_code = 0;
char g_err_msg[512] = "";

// The reference's profiling counters (src/profiling/Profiling.hpp:14-77):
// LUSTRINE_MAX_NUM_MEASUREMENTS slots holding the last measurement, cycles
// and seconds; its Unix build converts rdtsc cycles with a fixed 2.7 GHz.
constexpr int kProfSlots = 7;
constexpr double kProfHz = 2700000000.0;
double g_prof_dur[kProfSlots] = {0, 0, 0, 0, 0, 0, 0};
long long g_prof_cyc[kProfSlots] = {0, 0, 0, 0, 0, 0, 0};

void prof_store(const double* sec) {
    for (int i = 0; i < kProfSlots; ++i) {
        g_prof_dur[i] = sec[i];
        g_prof_cyc[i] = (long long)std::llround(sec[i] * kProfHz);
    }
}

// hipEvents destroyed with the set
struct EventSet {
    std::vector<hipEvent_t> e;
    ~EventSet() {
        for (hipEvent_t x : e) (void)hipEventDestroy(x);
    }
    hipError_t create(int n) {
        for (int i = 0; i < n; ++i) {
            hipEvent_t x;
            hipError_t r = hipEventCreate(&x);
            if (r != hipSuccess) return r;
            e.push_back(x);
        }
        return hipSuccess;
    }
    double seconds(int a, int b) const {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, e[a], e[b]) != hipSuccess) return 0.0;
        return (double)ms * 1e-3;
    }
};

int set_error(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err_msg, sizeof(g_err_msg), fmt, ap);
    va_end(ap);
    g_err_code = code;
    fprintf(stderr, "lustrine_gpu: %s\n", g_err_msg);
    return code;
}

}  // namespace

int lus::report_error(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err_msg, sizeof(g_err_msg), fmt, ap);
    va_end(ap);
    g_err_code = code;
    fprintf(stderr, "lustrine_gpu: %s\n", g_err_msg);
    return code;
}

namespace {

#define LUS_HIP(call)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess)                                                              \
            return set_error(LUS_EHIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                             __FILE__, __LINE__);                                          \
    } while (0)

// device allocation owned by a list, freed together
struct Arena {
    std::vector<void*> ptrs;
    ~Arena() { release(); }
    void release() {
        for (void* p : ptrs) (void)hipFree(p);
        ptrs.clear();
    }
    template <class T>
    hipError_t alloc(T** out, size_t count) {
        void* p = nullptr;
        hipError_t e = hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T));
        if (e == hipSuccess) ptrs.push_back(p);
        *out = static_cast<T*>(p);
        return e;
    }
};

std::vector<float4> to_f4(const float* xyz, int n) {
    std::vector<float4> v((size_t)std::max(n, 0));
    for (int i = 0; i < n; ++i)
        v[i] = xyz ? make_float4(xyz[3 * (size_t)i], xyz[3 * (size_t)i + 1], xyz[3 * (size_t)i + 2], 0.0f)
                   : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    return v;
}

void from_f4(const std::vector<float4>& v, float* xyz) {
    for (size_t i = 0; i < v.size(); ++i) {
        xyz[3 * i] = v[i].x;
        xyz[3 * i + 1] = v[i].y;
        xyz[3 * i + 2] = v[i].z;
    }
}

// host restatement of get_cell_id for range checks of user input
int host_cell(const LusGrid& g, float x, float y, float z) {
    volatile float px = x / g.cell_size;
    volatile float py = y / g.cell_size;
    volatile float pz = z / g.cell_size;
    return ((int)py) * g.gridX * g.gridZ + ((int)px) * g.gridZ + ((int)pz);
}

bool grid_ok(const LusGrid& g) {
    return g.cell_size > 0.0f && g.gridX > 0 && g.gridY > 0 && g.gridZ > 0 &&
           (long long)g.gridX * g.gridY * g.gridZ < (1LL << 30);
}

// every point must map to a valid cell: 0 <= coord and (int)(coord/cs) < dim
bool points_in_grid(const LusGrid& g, const float* xyz, int n) {
    const int dims[3] = {g.gridX, g.gridY, g.gridZ};
    for (int i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) {
            float c = xyz[3 * (size_t)i + a];
            if (!(c >= 0.0f)) return false;
            volatile float q = c / g.cell_size;
            if ((int)q >= dims[a]) return false;
        }
    return true;
}

}  // namespace

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
struct PbfState {
    bool ready = false;
    LusPbfParams p{};
    lus::PbfK k{};
    int n = 0, C = 0;
    Arena mem;
    float* stage = nullptr;  // 6 n floats: packed positions and velocities for downloads
    float4 *x = nullptr, *x2 = nullptr, *v = nullptr, *v2 = nullptr, *xa = nullptr, *xl = nullptr,
           *xp = nullptr;
    int *id = nullptr, *id2 = nullptr, *key = nullptr, *slot = nullptr, *tmp = nullptr,
        *key2 = nullptr, *counts = nullptr, *cs = nullptr, *tiles = nullptr, *maxc = nullptr;
    unsigned long long* masks = nullptr;  // neighbour sets: lists (flat) or row masks (tiled)
    int nmaxc = 0;  // iterations with a block-max table
    int nblk = 0;   // blocks of one particle-parallel launch
    bool stepped = false;
    // batch of scenes (nullptr scenes: one scene)
    lus::SceneBox* scenes = nullptr;
    int *sid = nullptr, *sid2 = nullptr;
    int nscenes = 0;
    // hipGraph replay of two captured steps
    bool use_graph = false;
    int parity = 0, graph_parity = 0;
    float graph_dt = 0.0f;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    // learned warm start and density statistics
    void* feat = nullptr;  // bf16 [n][32] network inputs
    int* maxd = nullptr;   // per-block max of C, then per-block sum of max(0, C)
    int* flag = nullptr;   // 1: rerun the plain step
    int* counters = nullptr;  // [0] measured steps, [1] reruns
    int* retest = nullptr;    // one int: re-test count of the last step (lus_pbf_retest_count)
    float* hist = nullptr;    // rings of e (max) and of mean max(0, C)
    bool learned = false, check = false;
    int learned_iters = 1;
    float bound = 0.01f;
    // mixed sand + fluid scene: phases ping-pong like id
    bool mixed = false;
    int *ph = nullptr, *ph2 = nullptr;
    int n_fluid = 0;
    // tiled neighbour passes (LDS staging of each tile's 9 stencil rows);
    // off by default: measured slower than one thread per slot on gfx950
    bool tiled = false;
    int zt = 16;           // cells per chunk along z
    int *tcount = nullptr, *toff = nullptr, *tscan = nullptr;
    int4* tl = nullptr;    // tiles
    int tgrid = 0;         // blocks of a tiled launch
    int lcap = 0, lcap_x = 0;  // staged float4 per array (XSPH stages two arrays)
    int nred = 0;          // block-reduction entries the last step wrote
    int fblock = lus::kBlock;  // threads per block of the flat neighbour passes
    bool wsbuild = true;       // flat list build with wave-staged candidates
    int lanes = 0;             // lanes per particle of the list walks (0: automatic)
    int sort_tile = 0;  // cells per block of the sort's prefix scan (0: by the cell count, pbf_sort_tile)
    int* scan_status = nullptr;  // two rows of the one-pass scan's tile sums (launch_scan_lookback)
    // the next step's cell keys, written by the XSPH pass for step length
    // keys_dt (NaN: none; the next step then runs the key kernel)
    bool fuse_keys = true;
    float keys_dt = std::numeric_limits<float>::quiet_NaN();
};

// the network of the learned step (device copies), kept across scene inits
struct NetState {
    bool ready = false;
    Arena mem;
    float* params = nullptr;  // fp32 [kMlpParams]
    void* packed = nullptr;   // bf16 B-operand images
    float* mu = nullptr;
    float* inv_sigma = nullptr;
    float out_scale = 1.0f;
};

// one plain step that also writes training rows (lus_pbf_record_step)
struct Record {
    float* feat;  // fp32 [rows][32]
    float* tgt;   // fp32 [rows][3]
    int stride, off;
};

struct SandState {
    bool ready = false;
    LusSandParams p{};
    lus::SandK k{};
    int n = 0, ns = 0, C = 0;
    int prev_attract = 0;
    Arena mem;
    float4 *x = nullptr, *x2 = nullptr, *v = nullptr, *v2 = nullptr, *xs = nullptr, *A = nullptr,
           *B = nullptr, *solid = nullptr;
    int *att = nullptr, *att2 = nullptr, *key = nullptr, *slot = nullptr, *tmp = nullptr,
        *perm = nullptr, *key2 = nullptr, *counts = nullptr, *cs = nullptr, *tiles = nullptr,
        *ocs = nullptr, *solid_ref = nullptr, *nbr_off = nullptr, *nbr_cnt = nullptr,
        *nbr_cnt1 = nullptr, *nbr_idx = nullptr, *ntiles = nullptr;
    long nbr_cap = 0;
    Arena nbr_mem;
    float4* final_iter = nullptr;
    int* qbuf = nullptr;  // one int: cell-box count
    int* err = nullptr;   // one int: a simulate_sand_v3 prediction left the grid
    float* stage = nullptr;  // 6 n floats: packed positions and velocities for downloads
    bool stepped = false;
    // fixed index-buffer capacity (lus_sand_set_list_capacity): the step
    // enqueues without waiting; *fail marks a step whose lists did not fit
    long fixed_cap = 0;
    int* fail = nullptr;
    // hipGraph replay of whole steps (fixed capacity only), one graph per
    // launch configuration: the step's arguments and buffer assignment
    struct Graph {
        float dt = 0.0f;
        int attract = 0, blow = 0, prev_attract = 0, variant = 0;
        float player[3] = {0.0f, 0.0f, 0.0f};
        const void* x = nullptr;
        const void* A = nullptr;
        hipGraph_t g = nullptr;
        hipGraphExec_t exec = nullptr;
    };
    bool use_graph = false;
    std::vector<Graph> graphs;
};

struct LusGpu;
namespace {
void sand_drop_graphs(SandState& S);
int sand_check(LusGpu* c);
}

struct LusMlpTrainer;

struct LusGpu {
    int device = 0;
    hipStream_t stream = nullptr;
    bool profiling = false;  // record phase events (lus_profiling_enable)
    PbfState pbf;
    NetState net;
    SandState sand;
    lus::FluidState* fluid = nullptr;      // simulate_fluid state (lus_fluid.hip)
    std::vector<LusMlpTrainer*> trainers;  // live trainers, detached on destroy
};

int lus::ctx_device(const LusGpu* c) { return c->device; }
hipStream_t lus::ctx_stream(const LusGpu* c) { return c->stream; }
lus::FluidState*& lus::ctx_fluid(LusGpu* c) { return c->fluid; }

// A trainer belongs to one context.  Destroying the context first frees the
// trainer's device memory and detaches it (ctx = nullptr): later calls on it
// fail with LUS_ESTATE, lus_mlp_trainer_destroy still releases the handle.
struct LusMlpTrainer {
    LusGpu* ctx = nullptr;
    int device = 0;
    long cap = 0, rows = 0;
    int rec_calls = 0;
    Arena mem;
    float* feat = nullptr;  // [cap][32]
    float* tgt = nullptr;   // [cap][3]
    bool has_net = false;
    std::vector<float> net;
};

namespace {

// derived PBF constants; the CPU oracle derives the same values the same way
lus::PbfK pbf_constants(const LusPbfParams& p) {
    lus::PbfK k{};
    const double hfd = (double)p.h * (double)p.f;
    const float P6 = (float)(315.0 / (64.0 * 3.14 * std::pow(hfd, 9.0)));
    k.SPf = (float)(15.0 / (3.14 * std::pow(hfd, 6.0) * (double)p.f));
    volatile float h = p.h, f = p.f;
    k.h2 = h * h;
    k.hf = h * f;
    k.hf2 = k.hf * k.hf;
    k.f = f;
    k.f2 = f * f;
    volatile float mr0 = p.mass / p.rest_density;
    k.mr0sq = mr0 * mr0;
    k.c_rho = (float)((double)p.mass * (double)P6 / (double)p.rest_density);
    volatile float inv_r0 = 1.0f / p.rest_density;
    k.neg_inv_r0 = -inv_r0;
    volatile float dq = p.scorr_dq;
    volatile float dq2 = dq * dq;
    volatile float f2dq2 = k.f2 * dq2;
    volatile float tq = k.hf2 - f2dq2;
    volatile float tq2 = tq * tq;
    volatile float tq3 = tq2 * tq;
    k.inv_tq3 = 1.0f / tq3;
    k.scorr_k = p.scorr_k;
    const double iq = (double)k.inv_tq3;
    k.f_scorr12 = (float)(-(double)p.scorr_k * iq * iq * iq * iq);
    k.f_spf_neg_inv_r0 = (float)((double)k.SPf * (double)k.neg_inv_r0);
    k.f_mr0sq_spf2 = (float)((double)k.mr0sq * (double)k.SPf * (double)k.SPf);
    k.n_scorr = p.scorr_n;
    k.nph = 1;
    k.cW = p.c_xsph * P6;
    k.eps = p.epsilon;
    for (int a = 0; a < 3; ++a) {
        k.lo[a] = p.lo[a];
        k.hi[a] = p.hi[a];
        k.g[a] = p.gravity[a];
    }
    k.cs = p.cell_size;
    k.gX = p.gridX;
    k.gY = p.gridY;
    k.gZ = p.gridZ;
    k.gXZ = p.gridX * p.gridZ;
    k.exact = p.exact ? 1 : 0;
    return k;
}

// Tiles up to which the step's cell scan is one look-back pass.
constexpr int kLookbackTiles = 1024;

// The sort tile in use: the one set, else 4096 cells per scan block for grids
// of at least 2^19 cells (measured ~1 us faster per step at 1 M particles and
// in the 400 K mixed scene) and 2048 below (the same time at 32 K).
int pbf_sort_tile(const PbfState& P) {
    if (P.sort_tile) return P.sort_tile;
    return P.C + 1 >= (1 << 19) ? 4096 : lus::kScanTileCells;
}

// The look-back scan's two rows of tile sums for the scene's cell count, zeroed.
hipError_t pbf_alloc_scan(PbfState& P, Arena& m) {
    const size_t rows = (size_t)lus::scan_lookback_rows(P.C + 1);
    hipError_t e = m.alloc(&P.scan_status, 2 * rows);
    if (e != hipSuccess) return e;
    return hipMemset(P.scan_status, 0, 2 * rows * sizeof(int));
}

// Enqueues one PBF substep.  ev (nullable): LUS_PBF_NPHASES-indexed event log,
// ev_phase[k] receives the phase of event k.  rec (nullable): a plain step
// that also writes training rows.
//
// Plain step:   key, scan, scatter, permute, K x (lambda, delta-p), XSPH
// Learned step: key, scan, scatter, permute, features (neighbour sets),
//               network (x* + correction), L x (lambda, delta-p),
//               XSPH (+ density maximum), decide, then -- gated on the
//               decision -- restore x2, K x (lambda, delta-p), XSPH again:
//               the plain step from the same sorted state.
int pbf_enqueue(LusGpu* c, float dt, std::vector<hipEvent_t>* ev, std::vector<int>* ev_phase,
                const Record* rec = nullptr) {
    PbfState& P = c->pbf;
    hipStream_t s = c->stream;
    auto mark = [&](int phase) -> hipError_t {
        if (!ev) return hipSuccess;
        hipEvent_t e;
        hipError_t r = hipEventCreate(&e);
        if (r != hipSuccess) return r;
        ev->push_back(e);
        ev_phase->push_back(phase);
        // label the kernels timed since the previous mark (per-kernel timing)
        lus::KernelTimes& kt = lus::kernel_times();
        if (kt.on) kt.phase.resize(kt.start.size(), phase);
        return hipEventRecord(e, s);
    };
    // per-step constants: the scene-id buffers of a batch ping-pong like id
    lus::PbfK k = P.k;
    if (P.scenes) {
        k.scenes = P.scenes;
        k.sid = P.sid;
        k.sid_out = P.sid2;
    }
    if (P.mixed) {
        k.ph = P.ph;
        k.ph_out = P.ph2;
    }
    if (ev) LUS_HIP(mark(-1));
    // the keys, unless the previous step's XSPH pass wrote them for this dt;
    // keys it wrote for another dt left their histogram in counts: cleared
    // before the key kernel counts again
    if (!(P.keys_dt == dt)) {
        if (!std::isnan(P.keys_dt)) LUS_HIP(lus::launch_zero_ints(P.counts, P.C + 1, s));
        LUS_HIP(lus::launch_pbf_key(P.x, P.v, P.n, k, dt, P.key, P.slot, P.counts, s));
        LUS_HIP(mark(0));
    }
    P.keys_dt = std::numeric_limits<float>::quiet_NaN();
    // the cell scan: one pass with look-back up to kLookbackTiles tiles (the
    // histogram cleared by the scatter), else the tile passes (which clear it)
    const int sort_tile = pbf_sort_tile(P);
    const bool lookback = lus::scan_tile_count(P.C + 1, sort_tile) <= kLookbackTiles;
    if (lookback) {
        const size_t rows = (size_t)lus::scan_lookback_rows(P.C + 1);
        LUS_HIP(lus::launch_scan_lookback(P.counts, P.cs, P.C + 1, P.scan_status + P.parity * rows,
                                          P.scan_status + (P.parity ^ 1) * rows, sort_tile, s));
    } else {
        LUS_HIP(lus::launch_exclusive_scan(P.counts, P.cs, P.C + 1, P.tiles, 1, s, sort_tile));
    }
    LUS_HIP(mark(1));
    LUS_HIP(lus::launch_scatter(P.key, P.slot, P.cs, P.n, P.tmp, s, lookback ? P.counts : nullptr, P.C + 1));
    LUS_HIP(mark(2));
    LUS_HIP(lus::launch_pbf_permute(P.tmp, P.key, P.cs, P.x, P.v, P.id, P.n, k, dt, P.x2, P.xp,
                                    P.id2, P.key2, nullptr, s));
    LUS_HIP(mark(3));
    // tiles of the step (cell columns cut into chunks of P.zt cells), or flat launches
    lus::TileArgs ta, tx;
    if (P.tiled && !P.mixed) {
        LUS_HIP(lus::launch_tiles(P.cs, P.key2, k, P.zt, P.tcount, P.toff, P.tscan, P.tl, s));
        LUS_HIP(mark(11));
        ta.tiles = tx.tiles = P.tl;
        ta.ntiles = tx.ntiles = P.toff + lus::tile_chunks(k, P.zt);
        ta.grid = tx.grid = P.tgrid;
        ta.lcap = P.lcap;
        tx.lcap = P.lcap_x;
    }
    ta.block = tx.block = P.fblock;
    ta.stage = P.wsbuild ? 1 : 0;
    // small fused plain scenes: several lanes per particle in the list walks
    if (!ta.tiles && !P.mixed && !P.learned && !rec && !k.exact &&
        (P.lanes == lus::kSubLanes || (P.lanes == 0 && P.n <= lus::kSubMaxN)))
        ta.sub = tx.sub = lus::kSubLanes;
    // block-reduction entries the passes write
    const int nred = ta.tiles ? P.tgrid : lus::grid_for(P.n, P.fblock);
    const int K = P.p.iterations;
    const bool learned = P.learned && !rec;
    if (P.mixed) {
        // mixed sand + fluid: lambda (fluid), correction (fluid / sand), XSPH (fluid)
        for (int it = 0; it < K; ++it) {
            LUS_HIP(lus::launch_mix_lambda(it == 0 ? P.xp : P.xa, P.key2, P.cs, P.n, k, P.xl, P.masks, P.xp,
                                           it == 0 ? 1 : 0, ta.stage, s));
            LUS_HIP(mark(it == 0 ? 7 : 4));
            LUS_HIP(lus::launch_mix_correct(P.xl, P.key2, P.cs, P.n, k, P.xa, P.x2, it == K - 1 ? P.v2 : nullptr,
                                            dt, P.masks, P.xp, s));
            LUS_HIP(mark(5));
        }
        if (K == 0) {
            LUS_HIP(lus::launch_mix_lambda(P.xp, P.key2, P.cs, P.n, k, P.xl, P.masks, P.xp, 1, ta.stage, s));
            LUS_HIP(lus::launch_pbf_velocity(P.xp, P.x2, P.n, dt, P.v2, s));
        }
        LUS_HIP(lus::launch_mix_xsph(K > 0 ? P.xa : P.xp, P.v2, P.key2, P.cs, P.n, k, P.v, P.x2, P.masks, P.xp,
                                     P.check ? P.maxd : nullptr, s));
        LUS_HIP(mark(6));
        if (P.check) {
            LUS_HIP(lus::launch_pbf_decide(P.maxd, lus::grid_for(P.n), std::max(P.n_fluid, 1), P.bound, 0, P.flag, P.counters,
                                           P.hist, LUS_DENSITY_HISTORY, s));
            LUS_HIP(mark(10));
        }
        P.nred = lus::grid_for(P.n);
        std::swap(P.x, P.x2);
        std::swap(P.id, P.id2);
        std::swap(P.ph, P.ph2);
        P.parity ^= 1;
        P.stepped = true;
        return LUS_OK;
    }
    // The plain iterations: xp (predicted) -> lambda -> xl -> delta-p -> xa ->
    // lambda -> xl ...; the last delta-p pass also writes v2 = (x* - x) / dt.
    // built: the neighbour masks exist already (features pass).
    auto plain_iterations = [&](bool built, const int* gate) -> int {
        for (int it = 0; it < K; ++it) {
            const float4* cur = it == 0 ? P.xp : P.xa;
            const int build = (it == 0 && !built) ? 1 : 0;
            LUS_HIP(lus::launch_pbf_lambda(cur, P.key2, P.cs, P.n, k, P.xl, P.maxc + (size_t)it * P.nblk,
                                           P.masks, P.xp, build, gate, ta, s));
            LUS_HIP(mark(it == 0 ? 7 : 4));
            const bool last = (it == K - 1);
            LUS_HIP(lus::launch_pbf_deltap(P.xl, P.key2, P.cs, P.n, k, P.xa, P.x2,
                                           last ? P.v2 : nullptr, dt, P.masks, P.xp, gate, ta, s));
            LUS_HIP(mark(5));
        }
        return LUS_OK;
    };
    if (learned) {
        const NetState& N = c->net;
        LUS_HIP(lus::launch_pbf_features(P.key2, P.cs, P.n, k, P.masks, P.xp, P.x2, N.mu, N.inv_sigma,
                                         P.feat, nullptr, 1, 0, ta, s));
        LUS_HIP(mark(8));
        LUS_HIP(lus::launch_mlp_forward(P.feat, P.n, N.params, N.packed, N.out_scale, nullptr, P.xp,
                                        P.xa, k, s));
        LUS_HIP(mark(9));
        const int L = P.learned_iters;
        for (int it = 0; it < L; ++it) {
            LUS_HIP(lus::launch_pbf_lambda(P.xa, P.key2, P.cs, P.n, k, P.xl, P.maxc, P.masks, P.xp, 0,
                                           nullptr, ta, s));
            LUS_HIP(mark(4));
            LUS_HIP(lus::launch_pbf_deltap(P.xl, P.key2, P.cs, P.n, k, P.xa, P.x2,
                                           it == L - 1 ? P.v2 : nullptr, dt, P.masks, P.xp, nullptr, ta, s));
            LUS_HIP(mark(5));
        }
        if (L == 0) LUS_HIP(lus::launch_pbf_velocity(P.xa, P.x2, P.n, dt, P.v2, s));
        LUS_HIP(lus::launch_pbf_xsph(P.xa, P.v2, P.key2, P.cs, P.n, k, P.v, P.x2, P.masks, P.xp, P.maxd,
                                     nullptr, tx, s));
        LUS_HIP(mark(6));
        LUS_HIP(lus::launch_pbf_decide(P.maxd, nred, P.n, P.bound, 1, P.flag, P.counters, P.hist,
                                       LUS_DENSITY_HISTORY, s));
        LUS_HIP(mark(10));
        // The plain step, run only when the decision asks for it (K >= 1).
        // x*, the neighbour masks and the overflow flags in key2 are what the
        // plain step's first lambda pass builds; only x2 (old positions, the
        // velocity reference) was overwritten by the learned commit.
        LUS_HIP(lus::launch_pbf_restore_old(P.tmp, P.key, P.cs, P.x, P.n, P.x2, P.flag, s));
        LUS_HIP(mark(3));
        if (int r = plain_iterations(true, P.flag)) return r;
        LUS_HIP(lus::launch_pbf_xsph(P.xa, P.v2, P.key2, P.cs, P.n, k, P.v, P.x2, P.masks, P.xp, P.maxd,
                                     P.flag, tx, s));
        LUS_HIP(mark(6));
        LUS_HIP(lus::launch_pbf_decide(P.maxd, nred, P.n, P.bound, 2, P.flag, P.counters, P.hist,
                                       LUS_DENSITY_HISTORY, s));
        LUS_HIP(mark(10));
    } else {
        if (rec) {
            LUS_HIP(lus::launch_pbf_features(P.key2, P.cs, P.n, k, P.masks, P.xp, P.x2, nullptr, nullptr,
                                             nullptr, rec->feat, rec->stride, rec->off, ta, s));
            LUS_HIP(mark(8));
        }
        if (int r = plain_iterations(rec != nullptr, nullptr)) return r;
        if (rec) LUS_HIP(lus::launch_pbf_record_target(P.xa, P.xp, P.n, rec->stride, rec->off, rec->tgt, s));
        const float4* fin = K > 0 ? P.xa : P.xp;
        if (K == 0) {
            // neighbour sets for XSPH (the lambda values are not used)
            if (!rec)
                LUS_HIP(lus::launch_pbf_lambda(P.xp, P.key2, P.cs, P.n, k, P.xl, P.maxc, P.masks, P.xp, 1,
                                               nullptr, ta, s));
            LUS_HIP(lus::launch_pbf_velocity(P.xp, P.x2, P.n, dt, P.v2, s));
        }
        // the next step's keys with this step's dt (its key kernel is then
        // skipped when it runs with the same dt)
        lus::NextKeys nk;
        const bool keys_out = P.fuse_keys && !tx.tiles;
        if (keys_out) {
            nk.key = P.key;
            nk.slot = P.slot;
            nk.counts = P.counts;
            nk.dt = dt;
        }
        LUS_HIP(lus::launch_pbf_xsph(fin, P.v2, P.key2, P.cs, P.n, k, P.v, P.x2, P.masks, P.xp,
                                     P.check ? P.maxd : nullptr, nullptr, tx, s, keys_out ? &nk : nullptr));
        if (keys_out) P.keys_dt = dt;
        LUS_HIP(mark(6));
        if (P.check) {
            LUS_HIP(lus::launch_pbf_decide(P.maxd, nred, P.n, P.bound, 0, P.flag, P.counters, P.hist,
                                           LUS_DENSITY_HISTORY, s));
            LUS_HIP(mark(10));
        }
    }
    // the last lambda pass's block maxima: ta.sub lanes per particle (the
    // building pass and the list walks alike)
    P.nred = ta.sub > 1 ? lus::grid_for(P.n * ta.sub) : nred;
    std::swap(P.x, P.x2);
    std::swap(P.id, P.id2);
    std::swap(P.sid, P.sid2);
    P.parity ^= 1;
    P.stepped = true;
    return LUS_OK;
}

int pbf_setup(LusGpu* c, const LusPbfParams* p, int n, const float* pos, const float* vel,
              const std::vector<lus::SceneBox>* scenes, const std::vector<int>* sid);

void pbf_drop_graph(PbfState& P) {
    if (P.graph_exec) (void)hipGraphExecDestroy(P.graph_exec);
    if (P.graph) (void)hipGraphDestroy(P.graph);
    P.graph_exec = nullptr;
    P.graph = nullptr;
}

// `steps` substeps; with graphs on, pairs of steps replay one captured graph
// (two steps restore the ping-pong buffer assignment the capture baked in)
int pbf_run(LusGpu* c, float dt, int steps) {
    PbfState& P = c->pbf;
    if (P.learned && (!c->net.ready || P.p.iterations < 1 || P.mixed))
        return set_error(LUS_ESTATE, "learned PBF step: needs a network, iterations >= 1 and a fluid scene");
    if (!P.use_graph) {
        for (int s = 0; s < steps; ++s) {
            int r = pbf_enqueue(c, dt, nullptr, nullptr);
            if (r) return r;
        }
        return LUS_OK;
    }
    int left = steps;
    while (left >= 2) {
        if (!P.graph_exec || P.graph_dt != dt || P.graph_parity != P.parity) {
            // a captured step that skips the key kernel relies on the step
            // before it: capture only when the keys at hand are exactly what
            // the captured steps leave behind (written for dt when the steps
            // write keys, none otherwise), so every replay sees the state the
            // capture saw
            const bool writes_keys = P.fuse_keys && !P.learned && !P.mixed && !P.tiled;
            if (writes_keys != (P.keys_dt == dt)) {
                int r = pbf_enqueue(c, dt, nullptr, nullptr);
                if (r) return r;
                --left;
                continue;
            }
            pbf_drop_graph(P);
            LUS_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            int r1 = pbf_enqueue(c, dt, nullptr, nullptr);
            int r2 = r1 ? r1 : pbf_enqueue(c, dt, nullptr, nullptr);
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(c->stream, &g);
            if (r1 || r2) {
                if (g) (void)hipGraphDestroy(g);
                return r1 ? r1 : r2;
            }
            LUS_HIP(e);
            P.graph = g;
            LUS_HIP(hipGraphInstantiate(&P.graph_exec, P.graph, nullptr, nullptr, 0));
            P.graph_dt = dt;
            P.graph_parity = P.parity;
        }
        LUS_HIP(hipGraphLaunch(P.graph_exec, c->stream));
        left -= 2;
    }
    if (left) return pbf_enqueue(c, dt, nullptr, nullptr);
    return LUS_OK;
}

}  // namespace

extern "C" {

int lus_gpu_last_error(char* buf, int n) {
    if (buf && n > 0) {
        strncpy(buf, g_err_msg, (size_t)n - 1);
        buf[n - 1] = '\0';
    }
    return g_err_code;
}

int lus_gpu_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int lus_gpu_create(LusGpu** out, int device) {
    if (!out) return set_error(LUS_EINVAL, "lus_gpu_create: out is NULL");
    *out = nullptr;
    int nd = lus_gpu_device_count();
    if (device < 0 || device >= nd)
        return set_error(LUS_EINVAL, "lus_gpu_create: device %d not available (%d visible)", device, nd);
    LUS_HIP(hipSetDevice(device));
    LusGpu* c = new LusGpu();
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return set_error(LUS_EHIP, "hipStreamCreateWithFlags: %s", hipGetErrorString(e));
    }
    *out = c;
    return LUS_OK;
}

void lus_gpu_destroy(LusGpu* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    pbf_drop_graph(c->pbf);
    sand_drop_graphs(c->sand);
    for (LusMlpTrainer* t : c->trainers) {  // detach: the handles stay valid for destroy
        t->mem.release();
        t->ctx = nullptr;
        t->rows = 0;
        t->cap = 0;
    }
    c->trainers.clear();
    c->pbf.mem.release();
    c->net.mem.release();
    c->sand.mem.release();
    c->sand.nbr_mem.release();
    lus::fluid_free(c->fluid);
    c->fluid = nullptr;
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int lus_gpu_stream_copy(LusGpu* c, size_t bytes, int reps, float* us) {
    if (!c || !us || reps <= 0 || bytes == 0 || bytes % 16 != 0)
        return set_error(LUS_EINVAL, "lus_gpu_stream_copy: bad arguments");
    LUS_HIP(hipSetDevice(c->device));
    Arena m;
    float4 *a, *b;
    const size_t n4 = bytes / 16;
    LUS_HIP(m.alloc(&a, n4));
    LUS_HIP(m.alloc(&b, n4));
    LUS_HIP(hipMemset(a, 0, bytes));
    LUS_HIP(hipDeviceSynchronize());
    LUS_HIP(lus::launch_stream_copy(a, b, n4, c->stream));  // warm
    hipEvent_t e0, e1;
    LUS_HIP(hipEventCreate(&e0));
    LUS_HIP(hipEventCreate(&e1));
    LUS_HIP(hipEventRecord(e0, c->stream));
    for (int r = 0; r < reps; ++r) LUS_HIP(lus::launch_stream_copy(a, b, n4, c->stream));
    LUS_HIP(hipEventRecord(e1, c->stream));
    LUS_HIP(hipEventSynchronize(e1));
    float ms = 0.0f;
    LUS_HIP(hipEventElapsedTime(&ms, e0, e1));
    *us = ms * 1e3f / (float)reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return LUS_OK;
}

int lus_gpu_sync(LusGpu* c) {
    if (!c) return set_error(LUS_EINVAL, "lus_gpu_sync: NULL context");
    LUS_HIP(hipStreamSynchronize(c->stream));
    return sand_check(c);
}

// ---------------------------------------------------------------------------
// building blocks
// ---------------------------------------------------------------------------
int lus_cell_ids(const LusGrid* g, int n, const float* pos, int* out) {
    if (!g || n < 0 || (n > 0 && (!pos || !out))) return set_error(LUS_EINVAL, "lus_cell_ids: bad arguments");
    if (!grid_ok(*g)) return set_error(LUS_EINVAL, "lus_cell_ids: bad grid");
    if (n == 0) return LUS_OK;
    Arena m;
    float4* d_pos;
    int* d_key;
    LUS_HIP(m.alloc(&d_pos, n));
    LUS_HIP(m.alloc(&d_key, n));
    std::vector<float4> h = to_f4(pos, n);
    LUS_HIP(hipMemcpy(d_pos, h.data(), sizeof(float4) * n, hipMemcpyHostToDevice));
    LUS_HIP(lus::launch_cell_keys(d_pos, n, *g, d_key, nullptr));
    LUS_HIP(hipMemcpy(out, d_key, sizeof(int) * n, hipMemcpyDeviceToHost));
    return LUS_OK;
}

int lus_counting_sort(int* counts, const int* keys, size_t n, size_t num_cells, int* sorted) {
    if (!counts || (n > 0 && (!keys || !sorted)) || num_cells == 0 || n >= (1u << 30) ||
        num_cells >= (1u << 30))
        return set_error(LUS_EINVAL, "lus_counting_sort: bad arguments");
    for (size_t i = 0; i < n; ++i)
        if (keys[i] < 0 || (size_t)keys[i] >= num_cells)
            return set_error(LUS_EINVAL, "lus_counting_sort: key %d at %zu outside [0, %zu)", keys[i], i,
                             num_cells);
    const int N = (int)n, C = (int)num_cells;
    Arena m;
    int *d_key, *d_slot, *d_counts, *d_cs, *d_tmp, *d_perm, *d_ks, *d_tiles;
    LUS_HIP(m.alloc(&d_key, N));
    LUS_HIP(m.alloc(&d_slot, N));
    LUS_HIP(m.alloc(&d_counts, C + 1));
    LUS_HIP(m.alloc(&d_cs, C + 1));
    LUS_HIP(m.alloc(&d_tmp, N));
    LUS_HIP(m.alloc(&d_perm, N));
    LUS_HIP(m.alloc(&d_ks, N));
    LUS_HIP(m.alloc(&d_tiles, lus::scan_tile_count(C + 1) + 1));
    if (N) LUS_HIP(hipMemcpy(d_key, keys, sizeof(int) * N, hipMemcpyHostToDevice));
    LUS_HIP(hipMemset(d_counts, 0, sizeof(int) * (C + 1)));
    if (N) LUS_HIP(lus::launch_count_keys(d_key, N, d_counts, d_slot, nullptr));
    LUS_HIP(lus::launch_exclusive_scan(d_counts, d_cs, C + 1, d_tiles, 0, nullptr));
    if (N) {
        LUS_HIP(lus::launch_scatter(d_key, d_slot, d_cs, N, d_tmp, nullptr));
        LUS_HIP(lus::launch_stable_rank(d_tmp, d_key, d_cs, N, d_perm, d_ks, nullptr));
        LUS_HIP(hipMemcpy(sorted, d_perm, sizeof(int) * N, hipMemcpyDeviceToHost));
    }
    LUS_HIP(hipMemcpy(counts, d_cs, sizeof(int) * (C + 1), hipMemcpyDeviceToHost));
    return LUS_OK;
}

namespace {
// stable cell sort of a point set on the device: sorted positions, keys,
// cell-start table and permutation (sorted slot -> input index)
struct SortedSet {
    float4* pos = nullptr;
    int *key = nullptr, *cs = nullptr, *perm = nullptr;
};

// solid positions and reference slots reordered by ascending slot
void slot_order(const float* xyz, const int* index, int n, std::vector<float>* pos,
                std::vector<int>* idx) {
    std::vector<int> o((size_t)n);
    for (int k = 0; k < n; ++k) o[k] = k;
    std::stable_sort(o.begin(), o.end(), [&](int a, int b) { return index[a] < index[b]; });
    pos->resize((size_t)3 * n);
    idx->resize((size_t)n);
    for (int k = 0; k < n; ++k) {
        for (int a = 0; a < 3; ++a) (*pos)[3 * (size_t)k + a] = xyz[3 * (size_t)o[k] + a];
        (*idx)[k] = index[o[k]];
    }
}

int sort_points(Arena& m, const LusGrid& g, const float* xyz, int n, SortedSet* out) {
    const int C = g.gridX * g.gridY * g.gridZ;
    float4 *d_pos, *d_sorted;
    int *d_key, *d_slot, *d_counts, *d_cs, *d_tmp, *d_perm, *d_ks, *d_tiles;
    LUS_HIP(m.alloc(&d_pos, n));
    LUS_HIP(m.alloc(&d_sorted, n));
    LUS_HIP(m.alloc(&d_key, n));
    LUS_HIP(m.alloc(&d_slot, n));
    LUS_HIP(m.alloc(&d_counts, C + 1));
    LUS_HIP(m.alloc(&d_cs, C + 1));
    LUS_HIP(m.alloc(&d_tmp, n));
    LUS_HIP(m.alloc(&d_perm, n));
    LUS_HIP(m.alloc(&d_ks, n));
    LUS_HIP(m.alloc(&d_tiles, lus::scan_tile_count(C + 1) + 1));
    LUS_HIP(hipMemset(d_counts, 0, sizeof(int) * (C + 1)));
    if (n) {
        std::vector<float4> h = to_f4(xyz, n);
        LUS_HIP(hipMemcpy(d_pos, h.data(), sizeof(float4) * n, hipMemcpyHostToDevice));
        LUS_HIP(lus::launch_cell_keys(d_pos, n, g, d_key, nullptr));
        LUS_HIP(lus::launch_count_keys(d_key, n, d_counts, d_slot, nullptr));
    }
    LUS_HIP(lus::launch_exclusive_scan(d_counts, d_cs, C + 1, d_tiles, 0, nullptr));
    if (n) {
        LUS_HIP(lus::launch_scatter(d_key, d_slot, d_cs, n, d_tmp, nullptr));
        LUS_HIP(lus::launch_stable_rank(d_tmp, d_key, d_cs, n, d_perm, d_ks, nullptr));
        LUS_HIP(lus::launch_gather4(d_pos, d_perm, n, d_sorted, nullptr));
    }
    out->pos = d_sorted;
    out->key = d_ks;
    out->cs = d_cs;
    out->perm = d_perm;
    return LUS_OK;
}
}  // namespace

int lus_neighbors_fluid(const LusGrid* g, float kernel_radius, int n_sand, const float* sand_pos,
                        int sand_start, int n_solid, const float* solid_pos, const int* solid_index,
                        long* off, int* idx, long cap, long* total) {
    if (!g || n_sand < 0 || n_solid < 0 || (n_sand && !sand_pos) ||
        (n_solid && (!solid_pos || !solid_index)) || !off || !total)
        return set_error(LUS_EINVAL, "lus_neighbors_fluid: bad arguments");
    if (!grid_ok(*g)) return set_error(LUS_EINVAL, "lus_neighbors_fluid: bad grid");
    if (!points_in_grid(*g, sand_pos, n_sand) || !points_in_grid(*g, solid_pos, n_solid))
        return set_error(LUS_EINVAL, "lus_neighbors_fluid: a point lies outside the grid");
    Arena m;
    SortedSet S, O;
    int r = sort_points(m, *g, sand_pos, n_sand, &S);
    if (r) return r;
    // solids enter their cells in ascending reference slot (Neighbors.cpp:380-384)
    std::vector<float> spos;
    std::vector<int> sidx;
    slot_order(solid_pos, solid_index, n_solid, &spos, &sidx);
    solid_pos = spos.data();
    solid_index = sidx.data();
    r = sort_points(m, *g, solid_pos, n_solid, &O);
    if (r) return r;
    // inverse permutation and sorted solid reference slots (host)
    std::vector<int> perm((size_t)n_sand), inv((size_t)n_sand);
    if (n_sand) LUS_HIP(hipMemcpy(perm.data(), S.perm, sizeof(int) * n_sand, hipMemcpyDeviceToHost));
    for (int p = 0; p < n_sand; ++p) inv[perm[p]] = p;
    std::vector<int> operm((size_t)n_solid), oref((size_t)n_solid);
    if (n_solid) LUS_HIP(hipMemcpy(operm.data(), O.perm, sizeof(int) * n_solid, hipMemcpyDeviceToHost));
    for (int p = 0; p < n_solid; ++p) oref[p] = solid_index[operm[p]];
    int *d_inv, *d_oref, *d_cnt, *d_idx;
    long* d_off;
    LUS_HIP(m.alloc(&d_inv, n_sand));
    LUS_HIP(m.alloc(&d_oref, n_solid));
    LUS_HIP(m.alloc(&d_cnt, n_sand));
    LUS_HIP(m.alloc(&d_off, n_sand + 1));
    if (n_sand) LUS_HIP(hipMemcpy(d_inv, inv.data(), sizeof(int) * n_sand, hipMemcpyHostToDevice));
    if (n_solid) LUS_HIP(hipMemcpy(d_oref, oref.data(), sizeof(int) * n_solid, hipMemcpyHostToDevice));
    volatile float h = kernel_radius;
    const float h2 = h * h;
    if (n_sand)
        LUS_HIP(lus::launch_fluid_nbr(S.pos, S.key, S.cs, S.perm, n_sand, sand_start,
                                      n_solid ? O.pos : nullptr, O.cs, d_oref, *g, h2, d_inv,
                                      nullptr, d_cnt, nullptr, 0, nullptr));
    std::vector<int> cnt((size_t)n_sand);
    if (n_sand) LUS_HIP(hipMemcpy(cnt.data(), d_cnt, sizeof(int) * n_sand, hipMemcpyDeviceToHost));
    off[0] = 0;
    for (int i = 0; i < n_sand; ++i) off[i + 1] = off[i] + cnt[i];
    *total = off[n_sand];
    if (*total > cap || (*total > 0 && !idx))
        return set_error(LUS_EINVAL, "lus_neighbors_fluid: capacity %ld < %ld entries", cap, *total);
    if (*total == 0) return LUS_OK;
    LUS_HIP(m.alloc(&d_idx, (size_t)*total));
    LUS_HIP(hipMemcpy(d_off, off, sizeof(long) * (n_sand + 1), hipMemcpyHostToDevice));
    LUS_HIP(lus::launch_fluid_nbr(S.pos, S.key, S.cs, S.perm, n_sand, sand_start,
                                  n_solid ? O.pos : nullptr, O.cs, d_oref, *g, h2, d_inv, d_off,
                                  d_cnt, d_idx, 1, nullptr));
    LUS_HIP(hipMemcpy(idx, d_idx, sizeof(int) * (size_t)*total, hipMemcpyDeviceToHost));
    return LUS_OK;
}

// ---------------------------------------------------------------------------
// PBF scene
// ---------------------------------------------------------------------------
int lus_pbf_init(LusGpu* c, const LusPbfParams* p, int n, const float* pos, const float* vel) {
    if (!c || !p || n <= 0 || !pos) return set_error(LUS_EINVAL, "lus_pbf_init: bad arguments");
    LusGrid g{p->cell_size, p->gridX, p->gridY, p->gridZ};
    if (!grid_ok(g)) return set_error(LUS_EINVAL, "lus_pbf_init: bad grid");
    if (!(p->h > 0.0f) || !(p->f > 0.0f) || p->cell_size < p->h || p->iterations < 0 ||
        p->iterations > 1024 || p->scorr_n < 1 || p->scorr_n > 16 || !(p->rest_density > 0.0f))
        return set_error(LUS_EINVAL, "lus_pbf_init: bad solver constants");
    // the clamp box must map into the grid, so every predicted position has a cell
    float corners[6] = {p->lo[0], p->lo[1], p->lo[2], p->hi[0], p->hi[1], p->hi[2]};
    if (!points_in_grid(g, corners, 2) || !(p->lo[0] <= p->hi[0]) || !(p->lo[1] <= p->hi[1]) ||
        !(p->lo[2] <= p->hi[2]))
        return set_error(LUS_EINVAL, "lus_pbf_init: clamp box outside the grid");
    return pbf_setup(c, p, n, pos, vel, nullptr, nullptr);
}

int lus_pbf_batch_init(LusGpu* c, const LusPbfParams* p, int nscenes, const int* counts,
                       const float* boxes, const float* pos, const float* vel) {
    if (!c || !p || nscenes <= 0 || !counts || !boxes || !pos)
        return set_error(LUS_EINVAL, "lus_pbf_batch_init: bad arguments");
    if (!(p->h > 0.0f) || !(p->f > 0.0f) || p->cell_size < p->h || p->iterations < 0 ||
        p->iterations > 1024 || p->scorr_n < 1 || p->scorr_n > 16 || !(p->rest_density > 0.0f) ||
        !(p->cell_size > 0.0f))
        return set_error(LUS_EINVAL, "lus_pbf_batch_init: bad solver constants");
    // scene grids stacked along y with one empty cell layer between scenes
    std::vector<lus::SceneBox> tab((size_t)nscenes);
    std::vector<int> sid;
    long long total = 0;
    int gX = 0, gY = 0, gZ = 0;
    for (int s = 0; s < nscenes; ++s) {
        if (counts[s] <= 0) return set_error(LUS_EINVAL, "lus_pbf_batch_init: scene %d is empty", s);
        total += counts[s];
        int dims[3];
        for (int a = 0; a < 3; ++a) {
            volatile float q = boxes[3 * s + a] / p->cell_size;
            dims[a] = (int)q + 1;
            tab[s].lo[a] = p->lo[a];
            tab[s].hi[a] = boxes[3 * s + a] - p->lo[a];
        }
        LusGrid gs{p->cell_size, dims[0], dims[1], dims[2]};
        float corners[6] = {tab[s].lo[0], tab[s].lo[1], tab[s].lo[2],
                            tab[s].hi[0], tab[s].hi[1], tab[s].hi[2]};
        if (!points_in_grid(gs, corners, 2) || !(tab[s].lo[0] <= tab[s].hi[0]) ||
            !(tab[s].lo[1] <= tab[s].hi[1]) || !(tab[s].lo[2] <= tab[s].hi[2]))
            return set_error(LUS_EINVAL, "lus_pbf_batch_init: scene %d clamp box outside its grid", s);
        tab[s].yoff = gY + (s ? 1 : 0);
        gY = tab[s].yoff + dims[1];
        gX = std::max(gX, dims[0]);
        gZ = std::max(gZ, dims[2]);
    }
    if (total >= (1LL << 30)) return set_error(LUS_EINVAL, "lus_pbf_batch_init: too many particles");
    sid.reserve((size_t)total);
    for (int s = 0; s < nscenes; ++s) sid.insert(sid.end(), (size_t)counts[s], s);
    LusPbfParams q = *p;
    q.gridX = gX;
    q.gridY = gY;
    q.gridZ = gZ;
    for (int a = 0; a < 3; ++a) {
        q.lo[a] = tab[0].lo[a];
        q.hi[a] = tab[0].hi[a];
    }
    LusGrid g{q.cell_size, gX, gY, gZ};
    if (!grid_ok(g)) return set_error(LUS_EINVAL, "lus_pbf_batch_init: batch grid too large");
    return pbf_setup(c, &q, (int)total, pos, vel, &tab, &sid);
}

int lus_mix_init(LusGpu* c, const LusPbfParams* p, const LusMixParams* mp, int n, const float* pos,
                 const float* vel, const int* phase) {
    if (!c || !p || !mp || n <= 0 || !pos || !phase) return set_error(LUS_EINVAL, "lus_mix_init: bad arguments");
    if (!(mp->diameter > 0.0f) || !(mp->collision >= 0.0f) || !(mp->friction >= 0.0f) || !(mp->mu_s >= 0.0f) ||
        !(mp->mu_k >= 0.0f))
        return set_error(LUS_EINVAL, "lus_mix_init: bad sand constants");
    int nf = 0;
    for (int i = 0; i < n; ++i) {
        if (phase[i] < 0 || phase[i] > 2) return set_error(LUS_EINVAL, "lus_mix_init: phase %d at %d", phase[i], i);
        nf += phase[i] == 0;
    }
    if (c->pbf.learned) return set_error(LUS_EINVAL, "lus_mix_init: learned mode is on");
    int r = lus_pbf_init(c, p, n, pos, vel);
    if (r) return r;
    PbfState& P = c->pbf;
    // cell columns split into fluid, sand and solid runs: three times the keys
    const long long C3 = 3ll * p->gridX * p->gridY * p->gridZ;
    if (C3 >= (1ll << 30)) return set_error(LUS_EINVAL, "lus_mix_init: grid too large (3 x %lld cells)", C3 / 3);
    P.C = (int)C3;
    P.k.nph = 3;
    LUS_HIP(P.mem.alloc(&P.counts, P.C + 1));
    LUS_HIP(P.mem.alloc(&P.cs, P.C + 1));
    LUS_HIP(P.mem.alloc(&P.tiles, lus::scan_tile_count(P.C + 1, lus::kScanTileMin) + 1));
    LUS_HIP(pbf_alloc_scan(P, P.mem));
    LUS_HIP(hipMemset(P.counts, 0, sizeof(int) * (P.C + 1)));
    LUS_HIP(P.mem.alloc(&P.ph, n));
    LUS_HIP(P.mem.alloc(&P.ph2, n));
    LUS_HIP(hipMemcpy(P.ph, phase, sizeof(int) * n, hipMemcpyHostToDevice));
    LUS_HIP(hipDeviceSynchronize());
    P.k.mix_d = mp->diameter;
    {
        const double far = (double)mp->diameter * (1.0 + 0x1p-19);
        float f2 = (float)(far * far);
        if ((double)f2 < far * far) f2 = std::nextafter(f2, INFINITY);
        P.k.mix_d2_far = f2;
    }
    P.k.mix_collision = mp->collision;
    P.k.mix_friction = mp->friction;
    P.k.mix_mu_s = mp->mu_s;
    P.k.mix_mu_k = mp->mu_k;
    P.mixed = true;
    P.n_fluid = nf;
    return LUS_OK;
}

}  // extern "C"

namespace {
int pbf_setup(LusGpu* c, const LusPbfParams* p, int n, const float* pos, const float* vel,
              const std::vector<lus::SceneBox>* scenes, const std::vector<int>* sid) {
    // the neighbour passes address particle data with 32-bit byte offsets
    if (n > lus::kMaxParticles)
        return set_error(LUS_EINVAL, "PBF scene: %d particles, at most %d", n, lus::kMaxParticles);
    LUS_HIP(hipSetDevice(c->device));
    PbfState& P = c->pbf;
    LUS_HIP(hipStreamSynchronize(c->stream));
    pbf_drop_graph(P);
    P.mem.release();
    P.stage = nullptr;
    P.ready = false;
    P.stepped = false;
    P.parity = 0;
    P.keys_dt = std::numeric_limits<float>::quiet_NaN();
    P.scenes = nullptr;
    P.sid = P.sid2 = nullptr;
    P.mixed = false;
    P.ph = P.ph2 = nullptr;
    P.n_fluid = n;
    P.nscenes = scenes ? (int)scenes->size() : 0;
    P.p = *p;
    P.k = pbf_constants(*p);
    P.n = n;
    P.C = p->gridX * p->gridY * p->gridZ;
    P.nmaxc = std::max(p->iterations, 1);
    Arena& m = P.mem;
    if (scenes) {
        LUS_HIP(m.alloc(&P.scenes, scenes->size()));
        LUS_HIP(m.alloc(&P.sid, n));
        LUS_HIP(m.alloc(&P.sid2, n));
        LUS_HIP(hipMemcpy(P.scenes, scenes->data(), sizeof(lus::SceneBox) * scenes->size(),
                          hipMemcpyHostToDevice));
        LUS_HIP(hipMemcpy(P.sid, sid->data(), sizeof(int) * n, hipMemcpyHostToDevice));
    }
    LUS_HIP(m.alloc(&P.x, n));
    LUS_HIP(m.alloc(&P.x2, n));
    LUS_HIP(m.alloc(&P.v, n));
    LUS_HIP(m.alloc(&P.v2, n));
    LUS_HIP(m.alloc(&P.xa, n));
    LUS_HIP(m.alloc(&P.xl, n));
    LUS_HIP(m.alloc(&P.xp, n));
    LUS_HIP(m.alloc(&P.masks, lus::nbr_words(n)));
    LUS_HIP(m.alloc(&P.id, n));
    LUS_HIP(m.alloc(&P.id2, n));
    LUS_HIP(m.alloc(&P.key, n));
    LUS_HIP(m.alloc(&P.slot, n));
    LUS_HIP(m.alloc(&P.tmp, n));
    LUS_HIP(m.alloc(&P.key2, n));
    LUS_HIP(m.alloc(&P.counts, P.C + 1));
    LUS_HIP(m.alloc(&P.cs, P.C + 1));
    LUS_HIP(m.alloc(&P.tiles, lus::scan_tile_count(P.C + 1, lus::kScanTileMin) + 1));
    LUS_HIP(pbf_alloc_scan(P, m));
    P.nblk = lus::grid_for(n, 64);  // block-reduction entries of the finest flat launch
    LUS_HIP(m.alloc(&P.maxc, (size_t)P.nmaxc * P.nblk));
    LUS_HIP(m.alloc(&P.maxd, 2 * (size_t)P.nblk));
    LUS_HIP(m.alloc(&P.flag, 1));
    LUS_HIP(m.alloc(&P.counters, 2));
    LUS_HIP(m.alloc(&P.retest, 1));
    LUS_HIP(m.alloc(&P.hist, 2 * LUS_DENSITY_HISTORY));
    LUS_HIP(hipMemset(P.counters, 0, 2 * sizeof(int)));
    LUS_HIP(hipMemset(P.flag, 0, sizeof(int)));
    {
        unsigned short* fb = nullptr;  // bf16 storage
        LUS_HIP(m.alloc(&fb, (size_t)n * lus::kFeat));
        P.feat = fb;
    }
    // tile buffers sized for the finest chunking (one cell per chunk), so the
    // chunk length can change without reallocation
    {
        const int nch = lus::tile_chunks(P.k, 1);
        LUS_HIP(m.alloc(&P.tcount, (size_t)nch + 1));
        LUS_HIP(m.alloc(&P.toff, (size_t)nch + 1));
        LUS_HIP(m.alloc(&P.tscan, (size_t)lus::scan_tile_count(nch + 1) + 1));
        LUS_HIP(m.alloc(&P.tl, (size_t)n / lus::kTileT + 1 + nch));
        LUS_HIP(hipMemset(P.tcount, 0, sizeof(int) * ((size_t)nch + 1)));
        int maxsh = 0;
        LUS_HIP(hipDeviceGetAttribute(&maxsh, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device));
        const int avail = maxsh - 2048;  // static shared memory of the tiled kernels
        P.lcap = std::min(2048, avail / 16);
        P.lcap_x = std::min(1536, avail / 32);
        P.tgrid = std::min(2048, lus::grid_for(n));
        if (P.lcap < 256 || P.lcap_x < 256) P.tiled = false;  // too little LDS to stage a tile
    }
    P.nred = lus::grid_for(n, P.fblock);
    std::vector<float4> hx = to_f4(pos, n), hv = to_f4(vel, n);
    std::vector<int> hid((size_t)n);
    for (int i = 0; i < n; ++i) hid[i] = i;
    LUS_HIP(hipMemcpy(P.x, hx.data(), sizeof(float4) * n, hipMemcpyHostToDevice));
    LUS_HIP(hipMemcpy(P.v, hv.data(), sizeof(float4) * n, hipMemcpyHostToDevice));
    LUS_HIP(hipMemcpy(P.id, hid.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    LUS_HIP(hipMemset(P.counts, 0, sizeof(int) * (P.C + 1)));
    std::vector<int> mc((size_t)P.nmaxc * P.nblk, INT32_MIN);
    LUS_HIP(hipMemcpy(P.maxc, mc.data(), sizeof(int) * mc.size(), hipMemcpyHostToDevice));
    // the context stream does not order against the null stream used above
    LUS_HIP(hipDeviceSynchronize());
    P.ready = true;
    return LUS_OK;
}
}  // namespace

extern "C" {

int lus_pbf_step(LusGpu* c, float dt, int steps) {
    if (!c || !c->pbf.ready) return set_error(LUS_ESTATE, "lus_pbf_step: no PBF scene");
    if (!(dt > 0.0f) || steps < 0) return set_error(LUS_EINVAL, "lus_pbf_step: bad dt/steps");
    LUS_HIP(hipSetDevice(c->device));
    if (!c->profiling || steps == 0) return pbf_run(c, dt, steps);
    // profiling: the last step runs with an event after every kernel
    int r = pbf_run(c, dt, steps - 1);
    if (r) return r;
    std::vector<hipEvent_t> ev;
    std::vector<int> ph;
    r = pbf_enqueue(c, dt, &ev, &ph);
    if (!r && hipEventSynchronize(ev.back()) != hipSuccess) r = set_error(LUS_EHIP, "lus_pbf_step: event");
    if (!r) {
        // slots: 3 predict, sort and tiles; 4 the solver iterations (with the
        // list build, features and network); 6 XSPH and the rerun decision;
        // 0 and 2 the whole step; 1 (Bullet) and 5 (boundary, fused) 0
        double sec[kProfSlots] = {0, 0, 0, 0, 0, 0, 0};
        for (size_t q = 1; q < ev.size(); ++q) {
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, ev[q - 1], ev[q]);
            const double t = (double)ms * 1e-3;
            const int p = ph[q];
            const int slot = (p <= 3 || p == 11) ? 3 : (p == 6 || p == 10) ? 6 : 4;
            sec[slot] += t;
            sec[0] += t;
        }
        sec[2] = sec[0];
        prof_store(sec);
    }
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    return r;
}

int lus_profiling_enable(LusGpu* c, int on) {
    if (!c) return set_error(LUS_EINVAL, "lus_profiling_enable: no context");
    c->profiling = on != 0;
    return LUS_OK;
}

long lus_get_num_observation(void) { return kProfSlots; }

long long lus_get_cycles(int index) { return (index >= 0 && index < kProfSlots) ? g_prof_cyc[index] : 0; }

double lus_get_duration(int index) { return (index >= 0 && index < kProfSlots) ? g_prof_dur[index] : 0.0; }

int lus_pbf_step_latency(LusGpu* c, float dt, int steps, float* ms) {
    if (!c || !c->pbf.ready || !ms || steps <= 0 || !(dt > 0.0f))
        return set_error(LUS_EINVAL, "lus_pbf_step_latency: bad arguments");
    LUS_HIP(hipSetDevice(c->device));
    const int per = c->pbf.use_graph ? 2 : 1;
    const int groups = (steps + per - 1) / per;
    std::vector<hipEvent_t> ev((size_t)groups + 1);
    for (auto& e : ev) LUS_HIP(hipEventCreate(&e));
    LUS_HIP(hipStreamSynchronize(c->stream));
    LUS_HIP(hipEventRecord(ev[0], c->stream));
    int done = 0;
    for (int g = 0; g < groups; ++g) {
        const int k = std::min(per, steps - done);
        int r = pbf_run(c, dt, k);
        if (r) return r;
        LUS_HIP(hipEventRecord(ev[g + 1], c->stream));
        done += k;
    }
    LUS_HIP(hipEventSynchronize(ev[groups]));
    done = 0;
    for (int g = 0; g < groups; ++g) {
        const int k = std::min(per, steps - done);
        float t = 0.0f;
        LUS_HIP(hipEventElapsedTime(&t, ev[g], ev[g + 1]));
        for (int q = 0; q < k; ++q) ms[done + q] = t / (float)k;
        done += k;
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    return LUS_OK;
}

int lus_pbf_set_graph(LusGpu* c, int on) {
    if (!c) return set_error(LUS_EINVAL, "lus_pbf_set_graph: NULL context");
    c->pbf.use_graph = on != 0;
    if (!on) {
        LUS_HIP(hipStreamSynchronize(c->stream));
        pbf_drop_graph(c->pbf);
    }
    return LUS_OK;
}

int lus_pbf_set_launch(LusGpu* c, int block) {
    if (!c) return set_error(LUS_EINVAL, "lus_pbf_set_launch: NULL context");
    if (block != 64 && block != 128 && block != 256)
        return set_error(LUS_EINVAL, "lus_pbf_set_launch: block must be 64, 128 or 256");
    LUS_HIP(hipSetDevice(c->device));
    LUS_HIP(hipStreamSynchronize(c->stream));
    pbf_drop_graph(c->pbf);
    c->pbf.fblock = block;
    return LUS_OK;
}

int lus_pbf_set_build(LusGpu* c, int staged) {
    if (!c) return set_error(LUS_EINVAL, "lus_pbf_set_build: NULL context");
    LUS_HIP(hipSetDevice(c->device));
    LUS_HIP(hipStreamSynchronize(c->stream));
    pbf_drop_graph(c->pbf);
    c->pbf.wsbuild = staged != 0;
    return LUS_OK;
}

int lus_pbf_set_lanes(LusGpu* c, int lanes) {
    if (!c) return set_error(LUS_EINVAL, "lus_pbf_set_lanes: NULL context");
    if (lanes != 0 && lanes != 1 && lanes != lus::kSubLanes)
        return set_error(LUS_EINVAL, "lus_pbf_set_lanes: lanes must be 0, 1 or %d", lus::kSubLanes);
    LUS_HIP(hipSetDevice(c->device));
    LUS_HIP(hipStreamSynchronize(c->stream));
    pbf_drop_graph(c->pbf);
    c->pbf.lanes = lanes;
    return LUS_OK;
}

int lus_pbf_set_sort_tile(LusGpu* c, int cells) {
    if (!c) return set_error(LUS_EINVAL, "lus_pbf_set_sort_tile: NULL context");
    if (cells != 0 && !lus::scan_tile_valid(cells))
        return set_error(LUS_EINVAL, "lus_pbf_set_sort_tile: cells must be 0, 1024, 2048 or 4096");
    LUS_HIP(hipSetDevice(c->device));
    LUS_HIP(hipStreamSynchronize(c->stream));
    pbf_drop_graph(c->pbf);
    c->pbf.sort_tile = cells;
    return LUS_OK;
}

int lus_pbf_set_key_fusion(LusGpu* c, int on) {
    if (!c) return set_error(LUS_EINVAL, "lus_pbf_set_key_fusion: NULL context");
    LUS_HIP(hipSetDevice(c->device));
    LUS_HIP(hipStreamSynchronize(c->stream));
    pbf_drop_graph(c->pbf);
    c->pbf.fuse_keys = on != 0;
    return LUS_OK;
}

int lus_pbf_set_tiling(LusGpu* c, int tiled, int zt) {
    if (!c) return set_error(LUS_EINVAL, "lus_pbf_set_tiling: NULL context");
    if (tiled && (zt < 1 || zt > lus::kTileZMax))
        return set_error(LUS_EINVAL, "lus_pbf_set_tiling: zt must be in [1, %d]", lus::kTileZMax);
    PbfState& P = c->pbf;
    if (tiled && P.ready && (P.lcap < 256 || P.lcap_x < 256))
        return set_error(LUS_EINVAL, "lus_pbf_set_tiling: not enough LDS per block on this device");
    LUS_HIP(hipSetDevice(c->device));
    LUS_HIP(hipStreamSynchronize(c->stream));
    pbf_drop_graph(P);
    P.tiled = tiled != 0;
    if (tiled) P.zt = zt;
    return LUS_OK;
}

int lus_pbf_download(LusGpu* c, int order, float* pos, float* vel, int* ids) {
    if (!c || !c->pbf.ready) return set_error(LUS_ESTATE, "lus_pbf_download: no PBF scene");
    PbfState& P = c->pbf;
    LUS_HIP(hipSetDevice(c->device));
    LUS_HIP(hipStreamSynchronize(c->stream));
    if (order != 0 && order != 1) return set_error(LUS_EINVAL, "lus_pbf_download: order must be 0 or 1");
    const int n = P.n;
    hipStream_t s = c->stream;
    // vec3 rows packed (and for order 1 put back in input order by id) on
    // the device: each array is one copy into the caller's memory
    if ((pos || vel) && !P.stage) LUS_HIP(P.mem.alloc(&P.stage, 6 * (size_t)n));
    const int* by = order == 1 ? P.id : nullptr;
    if (pos) {
        LUS_HIP(lus::launch_pack3(P.x, by, n, P.stage, s));
        LUS_HIP(hipMemcpyAsync(pos, P.stage, sizeof(float) * 3 * n, hipMemcpyDeviceToHost, s));
    }
    if (vel) {
        LUS_HIP(lus::launch_pack3(P.v, by, n, P.stage + 3 * (size_t)n, s));
        LUS_HIP(hipMemcpyAsync(vel, P.stage + 3 * (size_t)n, sizeof(float) * 3 * n, hipMemcpyDeviceToHost, s));
    }
    if (ids && order == 0) LUS_HIP(hipMemcpyAsync(ids, P.id, sizeof(int) * n, hipMemcpyDeviceToHost, s));
    LUS_HIP(hipStreamSynchronize(s));
    if (ids && order == 1)
        for (int r = 0; r < n; ++r) ids[r] = r;
    return LUS_OK;
}

int lus_pbf_last_max_constraint(LusGpu* c, float* out) {
    if (!c || !c->pbf.ready || !out) return set_error(LUS_ESTATE, "lus_pbf_last_max_constraint: no scene");
    PbfState& P = c->pbf;
    LUS_HIP(hipStreamSynchronize(c->stream));
    int last = std::max(P.p.iterations - 1, 0);
    std::vector<int> blk((size_t)P.nred);  // entries the last step's launches wrote
    LUS_HIP(hipMemcpy(blk.data(), P.maxc + (size_t)last * P.nblk, sizeof(int) * P.nred,
                      hipMemcpyDeviceToHost));
    int m = *std::max_element(blk.begin(), blk.end());
    int b = m >= 0 ? m : (m ^ 0x7FFFFFFF);
    float f;
    memcpy(&f, &b, sizeof(f));
    *out = f;
    return LUS_OK;
}

int lus_pbf_bench(LusGpu* c, float dt, int warmup, int steps, float* ms, float* phase_ms,
                  int* phase_launches) {
    if (!c || !c->pbf.ready || !ms || warmup < 0 || steps <= 0)
        return set_error(LUS_EINVAL, "lus_pbf_bench: bad arguments");
    LUS_HIP(hipSetDevice(c->device));
    int r = lus_pbf_step(c, dt, warmup);
    if (r) return r;
    LUS_HIP(hipStreamSynchronize(c->stream));
    if (!phase_ms) {
        hipEvent_t e0, e1;
        LUS_HIP(hipEventCreate(&e0));
        LUS_HIP(hipEventCreate(&e1));
        LUS_HIP(hipEventRecord(e0, c->stream));
        r = lus_pbf_step(c, dt, steps);
        if (r) return r;
        LUS_HIP(hipEventRecord(e1, c->stream));
        LUS_HIP(hipEventSynchronize(e1));
        LUS_HIP(hipEventElapsedTime(ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    } else {
        *ms = 0.0f;
        for (int k = 0; k < LUS_PBF_NPHASES; ++k) {
            phase_ms[k] = 0.0f;
            if (phase_launches) phase_launches[k] = 0;
        }
        // All `steps` are enqueued before one synchronisation, so the host
        // stays ahead of the device as in the untimed run.  Every kernel is
        // timed by an event pair stamped at its own dispatch's start and end
        // (these agree with a kernel trace's durations; events recorded
        // between launches would add the dispatch gaps, 2-4 us per kernel);
        // the phase marks label the kernels and count the phases.
        lus::KernelTimes& kt = lus::kernel_times();
        std::vector<std::vector<hipEvent_t>> ev((size_t)steps);
        std::vector<std::vector<int>> ph((size_t)steps);
        kt.start.clear();
        kt.stop.clear();
        kt.phase.clear();
        kt.failed = false;
        for (int s = 0; s < steps && !r; ++s) {
            kt.on = true;
            r = pbf_enqueue(c, dt, &ev[s], &ph[s]);
            kt.on = false;
            kt.phase.resize(kt.start.size(), -1);
        }
        if (!r && kt.failed) r = set_error(LUS_EHIP, "lus_pbf_bench: kernel event");
        if (!r && hipStreamSynchronize(c->stream) != hipSuccess) r = set_error(LUS_EHIP, "lus_pbf_bench: sync");
        for (size_t q = 0; !r && q < kt.start.size(); ++q) {
            float t = 0.0f;
            if (hipEventElapsedTime(&t, kt.start[q], kt.stop[q]) != hipSuccess) {
                r = set_error(LUS_EHIP, "lus_pbf_bench: elapsed time");
                break;
            }
            if (kt.phase[q] >= 0) phase_ms[kt.phase[q]] += t;
            *ms += t;
        }
        for (int s = 0; !r && phase_launches && s < steps; ++s)
            for (size_t k = 1; k < ph[s].size(); ++k) phase_launches[ph[s][k]] += 1;
        for (auto& v : ev)
            for (hipEvent_t e : v) (void)hipEventDestroy(e);
        for (hipEvent_t e : kt.start) (void)hipEventDestroy(e);
        for (hipEvent_t e : kt.stop) (void)hipEventDestroy(e);
        kt.start.clear();
        kt.stop.clear();
        kt.phase.clear();
        if (r) return r;
    }
    return LUS_OK;
}

int lus_pbf_time_kernels(LusGpu* c, int reps, float* us) {
    if (!c || !c->pbf.ready || !us || reps <= 0)
        return set_error(LUS_EINVAL, "lus_pbf_time_kernels: bad arguments");
    PbfState& P = c->pbf;
    if (!P.stepped) return set_error(LUS_ESTATE, "lus_pbf_time_kernels: no step yet");
    LUS_HIP(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    hipEvent_t e0, e1;
    LUS_HIP(hipEventCreate(&e0));
    LUS_HIP(hipEventCreate(&e1));
    // the last step's tiles still describe the current cell order
    lus::TileArgs ta, tx;
    ta.block = tx.block = P.fblock;
    ta.stage = P.wsbuild ? 1 : 0;
    if (P.tiled) {
        ta.tiles = tx.tiles = P.tl;
        ta.ntiles = tx.ntiles = P.toff + lus::tile_chunks(P.k, P.zt);
        ta.grid = tx.grid = P.tgrid;
        ta.lcap = P.lcap;
        tx.lcap = P.lcap_x;
    }
    for (int kind = 0; kind < 4; ++kind) {
        LUS_HIP(hipStreamSynchronize(s));
        LUS_HIP(hipEventRecord(e0, s));
        for (int r = 0; r < reps; ++r) {
            // xl, xa and x2 are per-step scratch between steps; x, v, id hold the state
            if (kind == 0)
                LUS_HIP(lus::launch_pbf_lambda(P.xa, P.key2, P.cs, P.n, P.k, P.xl, P.maxc, P.masks, P.xp,
                                               0, nullptr, ta, s));
            else if (kind == 1)
                LUS_HIP(lus::launch_pbf_deltap(P.xl, P.key2, P.cs, P.n, P.k, P.xa, nullptr, nullptr,
                                               1.0f, P.masks, P.xp, nullptr, ta, s));
            else if (kind == 2)
                LUS_HIP(lus::launch_pbf_xsph(P.xa, P.v2, P.key2, P.cs, P.n, P.k, P.xl, P.x2, P.masks,
                                             P.xp, nullptr, nullptr, tx, s));
            else
                LUS_HIP(lus::launch_pbf_lambda(P.xp, P.key2, P.cs, P.n, P.k, P.xl, P.maxc, P.masks, P.xp,
                                               1, nullptr, ta, s));
        }
        LUS_HIP(hipEventRecord(e1, s));
        LUS_HIP(hipEventSynchronize(e1));
        float ms = 0.0f;
        LUS_HIP(hipEventElapsedTime(&ms, e0, e1));
        us[kind] = ms * 1e3f / (float)reps;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return LUS_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// learned warm start: network, mode, statistics, training
// ---------------------------------------------------------------------------
namespace {

constexpr int kChunk = 256;  // split-K chunk of the weight-gradient products

// device buffers of one forward/backward pass over up to B rows
struct MlpBufs {
    Arena mem;
    int B = 0, chunks = 0;
    unsigned short *Xb = nullptr, *H1 = nullptr, *H2 = nullptr, *d3 = nullptr, *dZ2 = nullptr,
                   *dZ1 = nullptr, *pbf = nullptr;  // bf16 storage
    float *Yb = nullptr, *Z1 = nullptr, *Z2 = nullptr, *Z3 = nullptr, *dH2 = nullptr, *dH1 = nullptr,
          *err = nullptr, *part = nullptr, *grad = nullptr, *acc = nullptr;
    hipError_t alloc(int rows) {
        B = rows;
        chunks = (rows + kChunk - 1) / kChunk;
        const size_t b = (size_t)rows;
        hipError_t e;
        if ((e = mem.alloc(&Xb, b * lus::kMlpIn)) || (e = mem.alloc(&H1, b * lus::kMlpHid)) ||
            (e = mem.alloc(&H2, b * lus::kMlpHid)) || (e = mem.alloc(&d3, b * lus::kMlpOut)) ||
            (e = mem.alloc(&dZ2, b * lus::kMlpHid)) || (e = mem.alloc(&dZ1, b * lus::kMlpHid)) ||
            (e = mem.alloc(&pbf, lus::kMlpParams)) || (e = mem.alloc(&Yb, b * lus::kMlpOut)) ||
            (e = mem.alloc(&Z1, b * lus::kMlpHid)) || (e = mem.alloc(&Z2, b * lus::kMlpHid)) ||
            (e = mem.alloc(&Z3, b * lus::kMlpOut)) || (e = mem.alloc(&dH2, b * lus::kMlpHid)) ||
            (e = mem.alloc(&dH1, b * lus::kMlpHid)) || (e = mem.alloc(&err, b)) ||
            (e = mem.alloc(&part, (size_t)chunks * (lus::kMlpHid + 1) * lus::kMlpHid)) ||
            (e = mem.alloc(&grad, lus::kMlpParams)) || (e = mem.alloc(&acc, 1)))
            return e;
        return hipSuccess;
    }
};

// Forward pass of the training kernels over `rows` rows of b.Xb / b.Yb with
// fp32 parameters p; err[r] = squared error of row r.  grad: also the
// backward pass, gradient of the mean loss into b.grad.
//   Z1 = X W1, H1 = bf16(silu(Z1 + b1)), Z2 = H1 W2, H2 = bf16(silu(Z2 + b2)),
//   o = H2 W3 + b3;  dW|db = [A^T; 1] dZ per layer (bf16 operands)
int mlp_pass(hipStream_t s, MlpBufs& b, const float* p, int rows, float inv_scale, bool grad) {
    using namespace lus;
    const unsigned short* W1 = b.pbf + kMlpW1;
    const unsigned short* W2 = b.pbf + kMlpW2;
    const unsigned short* W3 = b.pbf + kMlpW3;
    LUS_HIP(launch_to_bf16(p, kMlpParams, b.pbf, s));
    LUS_HIP(launch_gemm_bf16(b.Xb, kMlpIn, 1, W1, kMlpHid, 1, b.Z1, kMlpHid, rows, kMlpHid, kMlpIn, 32, 0, s));
    LUS_HIP(launch_bias_silu(b.Z1, p + kMlpB1, rows, kMlpHid, b.H1, s));
    LUS_HIP(launch_gemm_bf16(b.H1, kMlpHid, 1, W2, kMlpHid, 1, b.Z2, kMlpHid, rows, kMlpHid, kMlpHid, 64, 0, s));
    LUS_HIP(launch_bias_silu(b.Z2, p + kMlpB2, rows, kMlpHid, b.H2, s));
    LUS_HIP(launch_gemm_bf16(b.H2, kMlpHid, 1, W3, kMlpOut, 1, b.Z3, kMlpOut, rows, kMlpOut, kMlpHid, 64, 0, s));
    LUS_HIP(launch_out_loss(b.Z3, p + kMlpB3, b.Yb, rows, inv_scale, b.d3, b.err, s));
    if (!grad) return LUS_OK;
    const int ch = (rows + kChunk - 1) / kChunk;
    // layer 3: [H2^T; 1] d3 -> dW3|db3; dH2 = d3 W3^T
    LUS_HIP(launch_gemm_bf16(b.H2, 1, kMlpHid, b.d3, kMlpOut, 1, b.part, kMlpOut, kMlpHid + 1, kMlpOut, rows,
                             kChunk, 1, s));
    LUS_HIP(launch_sum_partials(b.part, ch, (kMlpHid + 1) * kMlpOut, b.grad + kMlpW3, s));
    LUS_HIP(launch_gemm_bf16(b.d3, kMlpOut, 1, W3, 1, kMlpOut, b.dH2, kMlpHid, rows, kMlpHid, kMlpOut, 32, 0, s));
    LUS_HIP(launch_silu_back(b.dH2, b.Z2, (long)rows * kMlpHid, b.dZ2, s));
    // layer 2
    LUS_HIP(launch_gemm_bf16(b.H1, 1, kMlpHid, b.dZ2, kMlpHid, 1, b.part, kMlpHid, kMlpHid + 1, kMlpHid, rows,
                             kChunk, 1, s));
    LUS_HIP(launch_sum_partials(b.part, ch, (kMlpHid + 1) * kMlpHid, b.grad + kMlpW2, s));
    LUS_HIP(launch_gemm_bf16(b.dZ2, kMlpHid, 1, W2, 1, kMlpHid, b.dH1, kMlpHid, rows, kMlpHid, kMlpHid, 64, 0, s));
    LUS_HIP(launch_silu_back(b.dH1, b.Z1, (long)rows * kMlpHid, b.dZ1, s));
    // layer 1
    LUS_HIP(launch_gemm_bf16(b.Xb, 1, kMlpIn, b.dZ1, kMlpHid, 1, b.part, kMlpHid, kMlpIn + 1, kMlpHid, rows,
                             kChunk, 1, s));
    LUS_HIP(launch_sum_partials(b.part, ch, (kMlpIn + 1) * kMlpHid, b.grad + kMlpW1, s));
    return LUS_OK;
}

bool net_layout_ok(int nfloats) { return nfloats == LUS_NET_FLOATS; }

}  // namespace

extern "C" {

int lus_pbf_set_network(LusGpu* c, const float* net, int nfloats) {
    if (!c || !net || !net_layout_ok(nfloats))
        return set_error(LUS_EINVAL, "lus_pbf_set_network: need %d floats", LUS_NET_FLOATS);
    LUS_HIP(hipSetDevice(c->device));
    LUS_HIP(hipStreamSynchronize(c->stream));
    pbf_drop_graph(c->pbf);  // out_scale is a kernel argument of a captured graph
    NetState& N = c->net;
    if (!N.params) {
        unsigned short* pk = nullptr;
        LUS_HIP(N.mem.alloc(&N.params, lus::kMlpParams));
        LUS_HIP(N.mem.alloc(&pk, lus::kMlpPackBytes / 2));
        LUS_HIP(N.mem.alloc(&N.mu, lus::kFeat));
        LUS_HIP(N.mem.alloc(&N.inv_sigma, lus::kFeat));
        N.packed = pk;
    }
    LUS_HIP(hipMemcpy(N.params, net + LUS_NET_PARAMS, sizeof(float) * lus::kMlpParams, hipMemcpyHostToDevice));
    LUS_HIP(hipMemcpy(N.mu, net + LUS_NET_MU, sizeof(float) * lus::kFeat, hipMemcpyHostToDevice));
    LUS_HIP(hipMemcpy(N.inv_sigma, net + LUS_NET_INV_SIGMA, sizeof(float) * lus::kFeat, hipMemcpyHostToDevice));
    LUS_HIP(lus::launch_mlp_pack(N.params, N.packed, c->stream));
    LUS_HIP(hipStreamSynchronize(c->stream));
    N.out_scale = net[LUS_NET_OUT_SCALE];
    N.ready = true;
    return LUS_OK;
}

int lus_pbf_set_mode(LusGpu* c, int learned, int learned_iterations, float bound, int check) {
    if (!c) return set_error(LUS_EINVAL, "lus_pbf_set_mode: NULL context");
    if (learned && !c->net.ready) return set_error(LUS_ESTATE, "lus_pbf_set_mode: no network uploaded");
    if (learned_iterations < 0 || learned_iterations > 64)
        return set_error(LUS_EINVAL, "lus_pbf_set_mode: learned_iterations must be in [0, 64]");
    PbfState& P = c->pbf;
    if (learned && P.ready && P.p.iterations < 1)
        return set_error(LUS_EINVAL, "lus_pbf_set_mode: the plain rerun needs iterations >= 1");
    if (learned && P.ready && P.mixed)
        return set_error(LUS_EINVAL, "lus_pbf_set_mode: no learned step for mixed sand + fluid scenes");
    LUS_HIP(hipSetDevice(c->device));
    LUS_HIP(hipStreamSynchronize(c->stream));
    pbf_drop_graph(P);
    P.learned = learned != 0;
    P.learned_iters = learned_iterations;
    P.bound = bound;
    P.check = check != 0;
    if (P.ready) {
        LUS_HIP(hipMemset(P.counters, 0, 2 * sizeof(int)));
        LUS_HIP(hipDeviceSynchronize());
    }
    return LUS_OK;
}

int lus_pbf_learned_stats(LusGpu* c, long long* steps, long long* reruns, float* errs, float* errs_mean,
                          int cap, int* nerrs) {
    if (!c || !c->pbf.ready) return set_error(LUS_ESTATE, "lus_pbf_learned_stats: no PBF scene");
    PbfState& P = c->pbf;
    LUS_HIP(hipSetDevice(c->device));
    LUS_HIP(hipStreamSynchronize(c->stream));
    int cnt[2];
    LUS_HIP(hipMemcpy(cnt, P.counters, sizeof(cnt), hipMemcpyDeviceToHost));
    if (steps) *steps = cnt[0];
    if (reruns) *reruns = cnt[1];
    int nh = std::min(cnt[0], LUS_DENSITY_HISTORY);
    nh = std::min(nh, std::max(cap, 0));
    if ((errs || errs_mean) && nh > 0) {
        std::vector<float> h(2 * LUS_DENSITY_HISTORY);
        LUS_HIP(hipMemcpy(h.data(), P.hist, sizeof(float) * h.size(), hipMemcpyDeviceToHost));
        for (int q = 0; q < nh; ++q) {
            const size_t r = (size_t)((unsigned)(cnt[0] - nh + q) % LUS_DENSITY_HISTORY);
            if (errs) errs[q] = h[r];
            if (errs_mean) errs_mean[q] = h[LUS_DENSITY_HISTORY + r];
        }
    }
    if (nerrs) *nerrs = (errs || errs_mean) ? nh : 0;
    return LUS_OK;
}

int lus_pbf_retest_count(LusGpu* c, long long* count) {
    if (!c || !c->pbf.ready || !count) return set_error(LUS_ESTATE, "lus_pbf_retest_count: no PBF scene");
    PbfState& P = c->pbf;
    if (!P.stepped || P.tiled) {
        *count = 0;  // no lists yet, or the tiled passes (row masks, no lists)
        return LUS_OK;
    }
    LUS_HIP(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    LUS_HIP(hipMemsetAsync(P.retest, 0, sizeof(int), s));
    LUS_HIP(lus::launch_pbf_count_retest(P.masks, P.n, P.retest, s));
    int v = 0;
    LUS_HIP(hipMemcpyAsync(&v, P.retest, sizeof(int), hipMemcpyDeviceToHost, s));
    LUS_HIP(hipStreamSynchronize(s));
    *count = v;
    return LUS_OK;
}

int lus_mlp_trainer_create(LusGpu* c, long capacity, LusMlpTrainer** out) {
    if (!c || !out || capacity <= 0 || capacity > (1L << 28))
        return set_error(LUS_EINVAL, "lus_mlp_trainer_create: bad arguments");
    *out = nullptr;
    LUS_HIP(hipSetDevice(c->device));
    LusMlpTrainer* t = new LusMlpTrainer();
    t->ctx = c;
    t->device = c->device;
    t->cap = capacity;
    hipError_t e = t->mem.alloc(&t->feat, (size_t)capacity * lus::kFeat);
    if (e == hipSuccess) e = t->mem.alloc(&t->tgt, (size_t)capacity * 3);
    if (e != hipSuccess) {
        delete t;
        return set_error(LUS_ENOMEM, "lus_mlp_trainer_create: %s", hipGetErrorString(e));
    }
    c->trainers.push_back(t);
    *out = t;
    return LUS_OK;
}

void lus_mlp_trainer_destroy(LusMlpTrainer* t) {
    if (!t) return;
    (void)hipSetDevice(t->device);
    if (t->ctx) {
        (void)hipStreamSynchronize(t->ctx->stream);
        auto& v = t->ctx->trainers;
        v.erase(std::remove(v.begin(), v.end(), t), v.end());
    }
    t->mem.release();
    delete t;
}

int lus_mlp_trainer_add(LusMlpTrainer* t, const float* feat, const float* target, long rows) {
    if (!t || !feat || !target || rows < 0 || t->rows + rows > t->cap)
        return set_error(LUS_EINVAL, "lus_mlp_trainer_add: bad arguments or capacity exceeded");
    if (!t->ctx) return set_error(LUS_ESTATE, "lus_mlp_trainer_add: the trainer's context was destroyed");
    LUS_HIP(hipSetDevice(t->ctx->device));
    LUS_HIP(hipStreamSynchronize(t->ctx->stream));
    LUS_HIP(hipMemcpy(t->feat + (size_t)t->rows * lus::kFeat, feat, sizeof(float) * lus::kFeat * rows,
                      hipMemcpyHostToDevice));
    LUS_HIP(hipMemcpy(t->tgt + (size_t)t->rows * 3, target, sizeof(float) * 3 * rows, hipMemcpyHostToDevice));
    t->rows += rows;
    return LUS_OK;
}

int lus_mlp_trainer_rows(LusMlpTrainer* t, long* rows) {
    if (!t || !rows) return set_error(LUS_EINVAL, "lus_mlp_trainer_rows: bad arguments");
    *rows = t->rows;
    return LUS_OK;
}

int lus_mlp_trainer_download(LusMlpTrainer* t, long first, long rows, float* feat, float* target) {
    if (!t || first < 0 || rows < 0 || first + rows > t->rows)
        return set_error(LUS_EINVAL, "lus_mlp_trainer_download: bad range");
    if (!t->ctx) return set_error(LUS_ESTATE, "lus_mlp_trainer_download: the trainer's context was destroyed");
    LUS_HIP(hipSetDevice(t->ctx->device));
    LUS_HIP(hipStreamSynchronize(t->ctx->stream));
    if (feat)
        LUS_HIP(hipMemcpy(feat, t->feat + (size_t)first * lus::kFeat, sizeof(float) * lus::kFeat * rows,
                          hipMemcpyDeviceToHost));
    if (target)
        LUS_HIP(hipMemcpy(target, t->tgt + (size_t)first * 3, sizeof(float) * 3 * rows, hipMemcpyDeviceToHost));
    return LUS_OK;
}

int lus_pbf_record_step(LusGpu* c, LusMlpTrainer* t, float dt, int stride, long* added) {
    if (!c || !t || t->ctx != c || stride < 1 || !(dt > 0.0f))
        return set_error(LUS_EINVAL, "lus_pbf_record_step: bad arguments");
    PbfState& P = c->pbf;
    if (!P.ready) return set_error(LUS_ESTATE, "lus_pbf_record_step: no PBF scene");
    if (P.p.iterations < 1) return set_error(LUS_EINVAL, "lus_pbf_record_step: needs iterations >= 1");
    const int off = t->rec_calls % stride;
    const int p0 = (stride - off % stride) % stride;
    const long count = p0 < P.n ? (long)(P.n - 1 - p0) / stride + 1 : 0;
    if (t->rows + count > t->cap)
        return set_error(LUS_EINVAL, "lus_pbf_record_step: trainer capacity exceeded (%ld + %ld > %ld)", t->rows,
                         count, t->cap);
    LUS_HIP(hipSetDevice(c->device));
    Record rec{t->feat + (size_t)t->rows * lus::kFeat, t->tgt + (size_t)t->rows * 3, stride, off};
    int r = pbf_enqueue(c, dt, nullptr, nullptr, &rec);
    if (r) return r;
    t->rows += count;
    t->rec_calls += 1;
    if (added) *added = count;
    return LUS_OK;
}

int lus_mlp_fit(LusMlpTrainer* t, const LusMlpFitParams* fp, float* train_loss, float* val_loss) {
    if (!t || !fp || fp->epochs < 1 || fp->epochs > 100000 || fp->batch < 32 || fp->batch % 32 ||
        fp->batch > (1 << 16) || !(fp->lr > 0.0f) || !(fp->lr_final >= 0.0f) || !(fp->val_fraction >= 0.0f) ||
        !(fp->val_fraction <= 0.5f))
        return set_error(LUS_EINVAL, "lus_mlp_fit: bad arguments");
    using namespace lus;
    if (!t->ctx) return set_error(LUS_ESTATE, "lus_mlp_fit: the trainer's context was destroyed");
    const long rows = t->rows;
    const long nval = (long)((double)rows * fp->val_fraction);
    const long ntr = rows - nval;
    const int B = fp->batch;
    if (ntr < B) return set_error(LUS_EINVAL, "lus_mlp_fit: %ld training rows < batch %d", ntr, B);
    LusGpu* c = t->ctx;
    hipStream_t s = c->stream;
    LUS_HIP(hipSetDevice(c->device));
    LUS_HIP(hipStreamSynchronize(s));
    Arena m;
    // normalisation of the inputs and scale of the targets, from the whole set
    double* st = nullptr;
    LUS_HIP(m.alloc(&st, 2 * kFeat + 6));
    LUS_HIP(launch_col_stats(t->feat, rows, kFeat, st, s));
    LUS_HIP(launch_col_stats(t->tgt, rows, 3, st + 2 * kFeat, s));
    std::vector<double> hs(2 * kFeat + 6);
    LUS_HIP(hipStreamSynchronize(s));
    LUS_HIP(hipMemcpy(hs.data(), st, sizeof(double) * hs.size(), hipMemcpyDeviceToHost));
    std::vector<float> mu(kFeat), inv(kFeat);
    for (int f = 0; f < kFeat; ++f) {
        const double mean = hs[2 * f] / (double)rows;
        const double var = std::max(hs[2 * f + 1] / (double)rows - mean * mean, 0.0);
        mu[f] = (float)mean;
        inv[f] = (float)(1.0 / std::max(std::sqrt(var), 1e-4));
    }
    double ss = 0.0;
    for (int q = 0; q < 3; ++q) ss += hs[2 * kFeat + 2 * q + 1];
    float out_scale = (float)std::sqrt(ss / (3.0 * (double)rows));
    if (!(out_scale > 0.0f)) out_scale = 1.0f;
    const float inv_scale = 1.0f / out_scale;
    // initial parameters: LeCun-uniform, last layer scaled by 0.1, zero biases
    std::mt19937_64 rng(fp->seed);
    std::vector<float> p((size_t)kMlpParams, 0.0f);
    auto fill = [&](int off, int fan_in, int count, float gain) {
        const float a = gain * (float)std::sqrt(3.0 / fan_in);
        std::uniform_real_distribution<float> u(-a, a);
        for (int q = 0; q < count; ++q) p[(size_t)off + q] = u(rng);
    };
    fill(kMlpW1, kMlpIn, kMlpIn * kMlpHid, 1.0f);
    fill(kMlpW2, kMlpHid, kMlpHid * kMlpHid, 1.0f);
    fill(kMlpW3, kMlpHid, kMlpHid * kMlpOut, 0.1f);
    // random split
    std::vector<int> order((size_t)rows);
    for (long r = 0; r < rows; ++r) order[r] = (int)r;
    std::shuffle(order.begin(), order.end(), rng);
    std::vector<int> val(order.begin(), order.begin() + nval), tr(order.begin() + nval, order.end());
    float *dp, *dm, *dv, *dmu, *dinv, *tl, *vl;
    int *dtr, *dval;
    LUS_HIP(m.alloc(&dp, kMlpParams));
    LUS_HIP(m.alloc(&dm, kMlpParams));
    LUS_HIP(m.alloc(&dv, kMlpParams));
    LUS_HIP(m.alloc(&dmu, kFeat));
    LUS_HIP(m.alloc(&dinv, kFeat));
    LUS_HIP(m.alloc(&tl, fp->epochs));
    LUS_HIP(m.alloc(&vl, fp->epochs));
    LUS_HIP(m.alloc(&dtr, ntr));
    LUS_HIP(m.alloc(&dval, std::max(nval, 1L)));
    LUS_HIP(hipMemcpy(dp, p.data(), sizeof(float) * kMlpParams, hipMemcpyHostToDevice));
    LUS_HIP(hipMemset(dm, 0, sizeof(float) * kMlpParams));
    LUS_HIP(hipMemset(dv, 0, sizeof(float) * kMlpParams));
    LUS_HIP(hipMemset(tl, 0, sizeof(float) * fp->epochs));
    LUS_HIP(hipMemset(vl, 0, sizeof(float) * fp->epochs));
    LUS_HIP(hipMemcpy(dmu, mu.data(), sizeof(float) * kFeat, hipMemcpyHostToDevice));
    LUS_HIP(hipMemcpy(dinv, inv.data(), sizeof(float) * kFeat, hipMemcpyHostToDevice));
    if (nval) LUS_HIP(hipMemcpy(dval, val.data(), sizeof(int) * nval, hipMemcpyHostToDevice));
    MlpBufs b;
    LUS_HIP(b.alloc(B));
    LUS_HIP(hipDeviceSynchronize());
    const long nb = ntr / B;
    const double total = (double)fp->epochs * (double)nb;
    long step = 0;
    for (int e = 0; e < fp->epochs; ++e) {
        std::shuffle(tr.begin(), tr.end(), rng);
        LUS_HIP(hipStreamSynchronize(s));
        LUS_HIP(hipMemcpy(dtr, tr.data(), sizeof(int) * ntr, hipMemcpyHostToDevice));
        for (long q = 0; q < nb; ++q) {
            LUS_HIP(launch_normalize_rows(t->feat, dtr + q * B, B, dmu, dinv, b.Xb, s));
            LUS_HIP(launch_gather_rows(t->tgt, dtr + q * B, B, 3, b.Yb, s));
            if (int r = mlp_pass(s, b, dp, B, inv_scale, true)) return r;
            ++step;
            const double cosw = 0.5 * (1.0 + std::cos(M_PI * (double)(step - 1) / total));
            const float lr = (float)(fp->lr_final + (fp->lr - fp->lr_final) * cosw);
            const float c1 = (float)(1.0 / (1.0 - std::pow(0.9, (double)step)));
            const float c2 = (float)(1.0 / (1.0 - std::pow(0.999, (double)step)));
            LUS_HIP(launch_adam(dp, b.grad, dm, dv, kMlpParams, lr, 0.9f, 0.999f, 1e-8f, c1, c2, s));
            LUS_HIP(launch_sum_into(b.err, B, tl + e, s));
        }
        for (long q = 0; q < nval; q += B) {
            const int rr = (int)std::min<long>(B, nval - q);
            LUS_HIP(launch_normalize_rows(t->feat, dval + q, rr, dmu, dinv, b.Xb, s));
            LUS_HIP(launch_gather_rows(t->tgt, dval + q, rr, 3, b.Yb, s));
            if (int r = mlp_pass(s, b, dp, rr, inv_scale, false)) return r;
            LUS_HIP(launch_sum_into(b.err, rr, vl + e, s));
        }
    }
    LUS_HIP(hipStreamSynchronize(s));
    std::vector<float> htl(fp->epochs), hvl(fp->epochs);
    LUS_HIP(hipMemcpy(htl.data(), tl, sizeof(float) * fp->epochs, hipMemcpyDeviceToHost));
    LUS_HIP(hipMemcpy(hvl.data(), vl, sizeof(float) * fp->epochs, hipMemcpyDeviceToHost));
    LUS_HIP(hipMemcpy(p.data(), dp, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost));
    for (int e = 0; e < fp->epochs; ++e) {
        if (train_loss) train_loss[e] = htl[e] / (float)(nb * B);
        if (val_loss) val_loss[e] = nval ? hvl[e] / (float)nval : NAN;
    }
    t->net.assign((size_t)LUS_NET_FLOATS, 0.0f);
    std::copy(mu.begin(), mu.end(), t->net.begin() + LUS_NET_MU);
    std::copy(inv.begin(), inv.end(), t->net.begin() + LUS_NET_INV_SIGMA);
    std::copy(p.begin(), p.end(), t->net.begin() + LUS_NET_PARAMS);
    t->net[LUS_NET_OUT_SCALE] = out_scale;
    t->has_net = true;
    return LUS_OK;
}

int lus_mlp_trainer_network(LusMlpTrainer* t, float* net) {
    if (!t || !net) return set_error(LUS_EINVAL, "lus_mlp_trainer_network: bad arguments");
    if (!t->has_net) return set_error(LUS_ESTATE, "lus_mlp_trainer_network: no fitted network");
    std::copy(t->net.begin(), t->net.end(), net);
    return LUS_OK;
}

int lus_mlp_loss_grad(LusGpu* c, const float* params, int rows, const float* X, const float* Y, float inv_scale,
                      float* loss, float* grad) {
    if (!c || !params || !X || !Y || rows < 1 || rows > (1 << 20))
        return set_error(LUS_EINVAL, "lus_mlp_loss_grad: bad arguments");
    using namespace lus;
    hipStream_t s = c->stream;
    LUS_HIP(hipSetDevice(c->device));
    LUS_HIP(hipStreamSynchronize(s));
    Arena m;
    float *dp, *dx;
    LUS_HIP(m.alloc(&dp, kMlpParams));
    LUS_HIP(m.alloc(&dx, (size_t)rows * kMlpIn));
    MlpBufs b;
    LUS_HIP(b.alloc(rows));
    LUS_HIP(hipMemcpy(dp, params, sizeof(float) * kMlpParams, hipMemcpyHostToDevice));
    LUS_HIP(hipMemcpy(dx, X, sizeof(float) * kMlpIn * rows, hipMemcpyHostToDevice));
    LUS_HIP(hipMemcpy(b.Yb, Y, sizeof(float) * kMlpOut * rows, hipMemcpyHostToDevice));
    LUS_HIP(hipMemset(b.acc, 0, sizeof(float)));
    LUS_HIP(hipDeviceSynchronize());
    LUS_HIP(launch_to_bf16(dx, (long)rows * kMlpIn, b.Xb, s));
    if (int r = mlp_pass(s, b, dp, rows, inv_scale, true)) return r;
    LUS_HIP(launch_sum_into(b.err, rows, b.acc, s));
    LUS_HIP(hipStreamSynchronize(s));
    float acc = 0.0f;
    LUS_HIP(hipMemcpy(&acc, b.acc, sizeof(float), hipMemcpyDeviceToHost));
    if (loss) *loss = acc / (float)rows;
    if (grad) LUS_HIP(hipMemcpy(grad, b.grad, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost));
    return LUS_OK;
}

int lus_mlp_infer(LusGpu* c, const float* net, int rows, const float* feat, float* out) {
    if (!c || !net || !feat || !out || rows < 1 || rows > (1 << 24))
        return set_error(LUS_EINVAL, "lus_mlp_infer: bad arguments");
    using namespace lus;
    hipStream_t s = c->stream;
    LUS_HIP(hipSetDevice(c->device));
    LUS_HIP(hipStreamSynchronize(s));
    Arena m;
    float *dp, *dmu, *dinv, *dx, *dout;
    unsigned short *pk, *xb;
    LUS_HIP(m.alloc(&dp, kMlpParams));
    LUS_HIP(m.alloc(&dmu, kFeat));
    LUS_HIP(m.alloc(&dinv, kFeat));
    LUS_HIP(m.alloc(&dx, (size_t)rows * kFeat));
    LUS_HIP(m.alloc(&dout, (size_t)rows * 3));
    LUS_HIP(m.alloc(&pk, kMlpPackBytes / 2));
    LUS_HIP(m.alloc(&xb, (size_t)rows * kFeat));
    LUS_HIP(hipMemcpy(dp, net + LUS_NET_PARAMS, sizeof(float) * kMlpParams, hipMemcpyHostToDevice));
    LUS_HIP(hipMemcpy(dmu, net + LUS_NET_MU, sizeof(float) * kFeat, hipMemcpyHostToDevice));
    LUS_HIP(hipMemcpy(dinv, net + LUS_NET_INV_SIGMA, sizeof(float) * kFeat, hipMemcpyHostToDevice));
    LUS_HIP(hipMemcpy(dx, feat, sizeof(float) * kFeat * rows, hipMemcpyHostToDevice));
    LUS_HIP(hipDeviceSynchronize());
    LUS_HIP(launch_mlp_pack(dp, pk, s));
    LUS_HIP(launch_normalize_rows(dx, nullptr, rows, dmu, dinv, xb, s));
    PbfK k{};
    LUS_HIP(launch_mlp_forward(xb, rows, dp, pk, net[LUS_NET_OUT_SCALE], dout, nullptr, nullptr, k, s));
    LUS_HIP(hipStreamSynchronize(s));
    LUS_HIP(hipMemcpy(out, dout, sizeof(float) * 3 * rows, hipMemcpyDeviceToHost));
    return LUS_OK;
}

// ---------------------------------------------------------------------------
// sand scene
// ---------------------------------------------------------------------------
int lus_sand_init(LusGpu* c, const LusSandParams* p, int n_sand, const float* pos, const float* vel,
                  const int* attracted, int n_solid, const float* solid_pos, const int* solid_index) {
    if (!c || !p || n_sand <= 0 || !pos || n_solid < 0 || (n_solid && (!solid_pos || !solid_index)))
        return set_error(LUS_EINVAL, "lus_sand_init: bad arguments");
    const LusGrid& g = p->grid;
    if (!grid_ok(g)) return set_error(LUS_EINVAL, "lus_sand_init: bad grid");
    if (p->substeps < 0 || p->substeps > 1024 || !(p->mass > 0.0f))
        return set_error(LUS_EINVAL, "lus_sand_init: bad constants");
    const float r = p->particleRadius;
    float corners[6] = {r, r, r, p->domain[0] - r, p->domain[1] - r, p->domain[2] - r};
    if (!points_in_grid(g, corners, 2))
        return set_error(LUS_EINVAL, "lus_sand_init: clamp box outside the grid");
    if (!points_in_grid(g, solid_pos, n_solid))
        return set_error(LUS_EINVAL, "lus_sand_init: a solid particle lies outside the grid");
    LUS_HIP(hipSetDevice(c->device));
    SandState& S = c->sand;
    LUS_HIP(hipStreamSynchronize(c->stream));
    sand_drop_graphs(S);
    S.mem.release();
    S.nbr_mem.release();
    // everything that pointed into the released memory
    S.nbr_idx = nullptr;
    S.nbr_cap = 0;
    S.fail = nullptr;
    S.fixed_cap = 0;
    S.use_graph = false;
    S.stage = nullptr;
    S.ready = false;
    S.stepped = false;
    S.p = *p;
    S.n = n_sand;
    S.ns = n_solid;
    S.C = g.gridX * g.gridY * g.gridZ;
    S.prev_attract = 0;
    lus::SandK& k = S.k;
    k = lus::SandK{};
    k.r = r;
    k.d = p->particleDiameter;
    {
        // (d (1 + 2^-19))^2 rounded up: sqrtf of anything above it exceeds d
        const double far = (double)k.d * (1.0 + 0x1p-19);
        float f2 = (float)(far * far);
        if ((double)f2 < far * far) f2 = std::nextafter(f2, INFINITY);
        k.d2_far = f2;
    }
    volatile float hh = p->kernelRadius;
    k.h2 = hh * hh;
    k.w = 1.0f / p->mass;
    k.mass_n = p->mass;
    k.mass_sum = p->mass + p->mass;
    k.cs = g.cell_size;
    for (int a = 0; a < 3; ++a) {
        k.hi[a] = p->domain[a] - r;
        k.g[a] = p->gravity[a];
    }
    k.collision = p->collision_coeff;
    k.friction = p->friction_coeff;
    k.mu_s = p->mu_s;
    k.mu_k = p->mu_k;
    k.attract_radius = p->attract_radius;
    k.blow_radius = p->blow_radius;
    k.attract_coeff = p->attract_coeff;
    k.blow_coeff = p->blow_coeff;
    k.gX = g.gridX;
    k.gY = g.gridY;
    k.gZ = g.gridZ;
    k.gXZ = g.gridX * g.gridZ;
    k.sym_limit = p->ptr_solid_start;
    k.variant = 0;
    k.bcoef = 0.9f;
    for (int a = 0; a < 3; ++a) k.dom[a] = p->domain[a];
    Arena& m = S.mem;
    const int n = n_sand;
    LUS_HIP(m.alloc(&S.x, n));
    LUS_HIP(m.alloc(&S.x2, n));
    LUS_HIP(m.alloc(&S.v, n));
    LUS_HIP(m.alloc(&S.v2, n));
    LUS_HIP(m.alloc(&S.xs, n));
    LUS_HIP(m.alloc(&S.A, n));
    LUS_HIP(m.alloc(&S.B, n));
    LUS_HIP(m.alloc(&S.att, n));
    LUS_HIP(m.alloc(&S.att2, n));
    LUS_HIP(m.alloc(&S.key, n));
    LUS_HIP(m.alloc(&S.slot, n));
    LUS_HIP(m.alloc(&S.tmp, n));
    LUS_HIP(m.alloc(&S.perm, n));
    LUS_HIP(m.alloc(&S.key2, n));
    LUS_HIP(m.alloc(&S.counts, S.C + 1));
    LUS_HIP(m.alloc(&S.cs, S.C + 1));
    LUS_HIP(m.alloc(&S.tiles, lus::scan_tile_count(S.C + 1) + 1));
    LUS_HIP(m.alloc(&S.nbr_off, n + 1));
    LUS_HIP(m.alloc(&S.nbr_cnt, n + 1));
    LUS_HIP(m.alloc(&S.nbr_cnt1, n));
    LUS_HIP(m.alloc(&S.ntiles, lus::scan_tile_count(n + 1) + 1));
    LUS_HIP(m.alloc(&S.qbuf, 1));
    LUS_HIP(m.alloc(&S.err, 1));
    std::vector<float4> hx = to_f4(pos, n), hv = to_f4(vel, n);
    std::vector<int> ha((size_t)n, 0);
    if (attracted) memcpy(ha.data(), attracted, sizeof(int) * n);
    LUS_HIP(hipMemcpy(S.x, hx.data(), sizeof(float4) * n, hipMemcpyHostToDevice));
    LUS_HIP(hipMemcpy(S.v, hv.data(), sizeof(float4) * n, hipMemcpyHostToDevice));
    LUS_HIP(hipMemcpy(S.att, ha.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    LUS_HIP(hipMemset(S.counts, 0, sizeof(int) * (S.C + 1)));
    LUS_HIP(hipMemset(S.nbr_cnt, 0, sizeof(int) * (n + 1)));
    // the out-of-grid flag is read by every no-wait step's check, whatever the
    // variant: it starts clear (a new allocation may hold an old context's data)
    LUS_HIP(hipMemset(S.err, 0, sizeof(int)));
    // static solids: stable cell sort once (uniform_grid_cells_static_saved,
    // src/neighbors/Neighbors.cpp:266-272, keeps ascending slot order per cell)
    SortedSet O;
    // the static cell lists hold solids in ascending reference slot
    // (Neighbors.cpp:266-272)
    std::vector<float> spos;
    std::vector<int> sidx;
    slot_order(solid_pos, solid_index, n_solid, &spos, &sidx);
    solid_pos = spos.data();
    solid_index = sidx.data();
    int rr = sort_points(m, g, solid_pos, n_solid, &O);
    if (rr) return rr;
    S.solid = n_solid ? O.pos : nullptr;
    S.ocs = O.cs;
    std::vector<int> operm((size_t)n_solid), oref((size_t)n_solid);
    if (n_solid) LUS_HIP(hipMemcpy(operm.data(), O.perm, sizeof(int) * n_solid, hipMemcpyDeviceToHost));
    for (int q = 0; q < n_solid; ++q) oref[q] = solid_index[operm[q]];
    LUS_HIP(m.alloc(&S.solid_ref, n_solid));
    if (n_solid) LUS_HIP(hipMemcpy(S.solid_ref, oref.data(), sizeof(int) * n_solid, hipMemcpyHostToDevice));
    LUS_HIP(hipDeviceSynchronize());
    S.ready = true;
    return LUS_OK;
}

namespace {

void sand_drop_graphs(SandState& S) {
    for (auto& g : S.graphs) {
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        if (g.g) (void)hipGraphDestroy(g.g);
    }
    S.graphs.clear();
}

// A step enqueued without waiting reported its failure on the device: the
// lists did not fit the fixed capacity, or (simulate_sand_v3) a predicted
// position left the grid.  Reported once, with the flags cleared.
int sand_check(LusGpu* c) {
    SandState& S = c->sand;
    // (the waiting step reports both itself)
    if (!S.ready || !S.fail || S.fixed_cap <= 0) return LUS_OK;
    int f[2] = {0, 0};
    LUS_HIP(hipMemcpy(&f[0], S.fail, sizeof(int), hipMemcpyDeviceToHost));
    if (S.err) LUS_HIP(hipMemcpy(&f[1], S.err, sizeof(int), hipMemcpyDeviceToHost));
    if (f[0]) LUS_HIP(hipMemset(S.fail, 0, sizeof(int)));
    if (f[1]) LUS_HIP(hipMemset(S.err, 0, sizeof(int)));
    if (f[1])
        return set_error(LUS_EINVAL, "lus_sand_step: a predicted position's cell lies outside the grid "
                                     "(simulate_sand_v3 does not clamp); the state is invalid");
    if (f[0])
        return set_error(LUS_ESTATE, "lus_sand_step: the neighbour lists exceeded the fixed capacity of %ld "
                                     "entries (lus_sand_set_list_capacity); the state is invalid",
                         S.fixed_cap);
    return LUS_OK;
}

// Enqueues one simulate_sand step.  Without a fixed capacity it waits once
// for the lists' total length (it sizes their buffer); with one it does not
// wait, and a step whose lists do not fit is reported by sand_check.  The
// host-side buffer swaps are applied by sand_after.
int sand_enqueue(LusGpu* c, float dt, const lus::SandK& k) {
    SandState& S = c->sand;
    hipStream_t s = c->stream;
    const int n = S.n;
    const bool fixed = S.fixed_cap > 0;
    EventSet pe;
    auto mark = [&](int q) -> hipError_t { return c->profiling ? hipEventRecord(pe.e[q], s) : hipSuccess; };
    if (c->profiling) LUS_HIP(pe.create(5));
    LUS_HIP(mark(0));
    if (k.variant == 2 && !fixed) LUS_HIP(lus::launch_zero_ints(S.err, 1, s));
    LUS_HIP(lus::launch_sand_predict_key(S.x, S.v, S.att, n, k, dt, S.err, S.xs, S.key, S.slot, S.counts,
                                         s));
    LUS_HIP(mark(1));
    LUS_HIP(lus::launch_exclusive_scan(S.counts, S.cs, S.C + 1, S.tiles, 1, s));
    LUS_HIP(lus::launch_scatter(S.key, S.slot, S.cs, n, S.tmp, s));
    LUS_HIP(lus::launch_sand_permute(S.tmp, S.key, S.cs, S.x, S.xs, S.v, S.att, n, S.x2, S.A, S.v2,
                                     S.att2, S.perm, S.key2, s));
    // neighbor lists: count, scan, fill
    LUS_HIP(lus::launch_sand_nbr(S.A, S.key2, S.cs, n, S.solid, S.ocs, k, nullptr, S.nbr_cnt,
                                 S.nbr_cnt1, nullptr, 0, s));
    LUS_HIP(lus::launch_exclusive_scan(S.nbr_cnt, S.nbr_off, n + 1, S.ntiles, 0, s));
    if (!fixed) {
        int total = 0, bad_key = 0;
        LUS_HIP(hipMemcpyAsync(&total, S.nbr_off + n, sizeof(int), hipMemcpyDeviceToHost, s));
        if (k.variant == 2) LUS_HIP(hipMemcpyAsync(&bad_key, S.err, sizeof(int), hipMemcpyDeviceToHost, s));
        LUS_HIP(hipStreamSynchronize(s));
        if (bad_key)
            return set_error(LUS_EINVAL,
                             "lus_sand_step: a predicted position's cell lies outside the grid (simulate_sand_v3 "
                             "does not clamp; the reference indexes outside its cell vector there)");
        if (total < 0) return set_error(LUS_EHIP, "lus_sand_step: neighbor count overflow");
        if ((long)total > S.nbr_cap || !S.nbr_idx) {
            S.nbr_mem.release();
            long cap = std::max<long>((long)total + total / 4, 64);
            LUS_HIP(S.nbr_mem.alloc(&S.nbr_idx, (size_t)cap));
            S.nbr_cap = cap;
        }
    }
    LUS_HIP(lus::launch_sand_nbr(S.A, S.key2, S.cs, n, S.solid, S.ocs, k, S.nbr_off, S.nbr_cnt,
                                 S.nbr_cnt1, S.nbr_idx, 1, s, S.nbr_cap, fixed ? S.fail : nullptr));
    LUS_HIP(mark(2));
    float4* A = S.A;
    float4* B = S.B;
    // simulate_sand_v3's substeps all read the same positions_tmp: one pass
    const int passes = k.variant == 2 ? (S.p.substeps > 0 ? 1 : 0) : S.p.substeps;
    for (int it = 0; it < passes; ++it) {
        LUS_HIP(lus::launch_sand_substep(A, S.x2, S.solid, S.nbr_off, S.nbr_idx, n, k, B, S.att2, s,
                                         fixed ? S.fail : nullptr));
        std::swap(A, B);
    }
    LUS_HIP(mark(3));
    LUS_HIP(lus::launch_sand_vupdate(A, S.x2, S.v2, n, dt, s));
    LUS_HIP(mark(4));
    if (c->profiling) {
        // slots of simulate_sand_v1 (src/Simulate.cpp:669-802): 0 the call,
        // 2 the solver, 3 the neighbour search, 4 the substeps, 6 the velocity
        // update; 1 (Bullet, on the host) and 5 (boundary, inside the
        // substep kernels here) are 0
        LUS_HIP(hipEventSynchronize(pe.e[4]));
        const double all = pe.seconds(0, 4);
        const double sec[kProfSlots] = {all, 0.0, all, pe.seconds(1, 2), pe.seconds(2, 3), 0.0, pe.seconds(3, 4)};
        prof_store(sec);
    }
    return LUS_OK;
}

// The buffer swaps and flags a step leaves behind (after sand_enqueue or a
// replayed graph of it)
void sand_after(SandState& S, const lus::SandK& k) {
    const int passes = k.variant == 2 ? (S.p.substeps > 0 ? 1 : 0) : S.p.substeps;
    if (passes & 1) std::swap(S.A, S.B);
    S.final_iter = S.A;
    std::swap(S.x, S.x2);
    std::swap(S.v, S.v2);
    std::swap(S.att, S.att2);
    if (k.variant != 2) S.prev_attract = k.attract;  // v3 keeps no previous flag
    S.stepped = true;
}

}  // namespace

int lus_sand_step(LusGpu* c, float dt, int attract_flag, int blow_flag, const float* player) {
    if (!c || !c->sand.ready) return set_error(LUS_ESTATE, "lus_sand_step: no sand scene");
    if (!(dt != 0.0f) || ((attract_flag || blow_flag) && !player))
        return set_error(LUS_EINVAL, "lus_sand_step: bad arguments");
    LUS_HIP(hipSetDevice(c->device));
    SandState& S = c->sand;
    lus::SandK k = S.k;
    k.attract = attract_flag ? 1 : 0;
    k.blow = blow_flag ? 1 : 0;
    k.prev_attract = S.prev_attract;
    // the player position matters only with a flag set (a graph key then
    // ignores it otherwise)
    for (int a = 0; a < 3; ++a) k.player[a] = (player && (k.attract || k.blow)) ? player[a] : 0.0f;
    if (k.variant == 2) {  // simulate_sand_v3: dt = clamp(dt, 0.001, 0.016) (src/Simulate.cpp:522)
        const float t = dt < 0.001f ? 0.001f : dt;
        dt = 0.016f < t ? 0.016f : t;
    }
    if (!S.use_graph || S.fixed_cap <= 0 || c->profiling) {
        if (int r = sand_enqueue(c, dt, k)) return r;
        sand_after(S, k);
        return LUS_OK;
    }
    SandState::Graph key;
    key.dt = dt;
    key.attract = k.attract;
    key.blow = k.blow;
    key.prev_attract = k.prev_attract;
    key.variant = k.variant;
    for (int a = 0; a < 3; ++a) key.player[a] = k.player[a];
    key.x = S.x;
    key.A = S.A;
    SandState::Graph* hit = nullptr;
    for (auto& g : S.graphs)
        if (g.dt == key.dt && g.attract == key.attract && g.blow == key.blow && g.prev_attract == key.prev_attract &&
            g.variant == key.variant && g.player[0] == key.player[0] && g.player[1] == key.player[1] &&
            g.player[2] == key.player[2] && g.x == key.x && g.A == key.A)
            hit = &g;
    if (!hit) {
        // a step's buffer assignments alternate, so a few graphs cover a run;
        // more configurations than that (a moving player with a flag set)
        // start over
        if (S.graphs.size() >= 8) sand_drop_graphs(S);
        LUS_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        const int r = sand_enqueue(c, dt, k);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(c->stream, &g);
        if (r) {
            if (g) (void)hipGraphDestroy(g);
            return r;
        }
        LUS_HIP(e);
        key.g = g;
        hipError_t ei = hipGraphInstantiate(&key.exec, g, nullptr, nullptr, 0);
        if (ei != hipSuccess) {
            (void)hipGraphDestroy(g);
            LUS_HIP(ei);
        }
        S.graphs.push_back(key);
        hit = &S.graphs.back();
    }
    LUS_HIP(hipGraphLaunch(hit->exec, c->stream));
    sand_after(S, k);
    return LUS_OK;
}

int lus_sand_set_list_capacity(LusGpu* c, long entries) {
    if (!c || !c->sand.ready) return set_error(LUS_ESTATE, "lus_sand_set_list_capacity: no sand scene");
    if (entries < 0 || entries > (long)INT_MAX)
        return set_error(LUS_EINVAL, "lus_sand_set_list_capacity: entries must be in [0, %d]", INT_MAX);
    LUS_HIP(hipSetDevice(c->device));
    SandState& S = c->sand;
    LUS_HIP(hipStreamSynchronize(c->stream));
    if (int r = sand_check(c)) return r;
    sand_drop_graphs(S);
    if (entries == 0) {
        S.fixed_cap = 0;
        S.use_graph = false;
        return LUS_OK;
    }
    S.nbr_mem.release();
    S.nbr_idx = nullptr;
    S.nbr_cap = 0;
    LUS_HIP(S.nbr_mem.alloc(&S.nbr_idx, (size_t)entries));
    S.nbr_cap = entries;
    if (!S.fail) LUS_HIP(S.mem.alloc(&S.fail, 1));
    LUS_HIP(hipMemset(S.fail, 0, sizeof(int)));
    S.fixed_cap = entries;
    return LUS_OK;
}

int lus_sand_set_graph(LusGpu* c, int on) {
    if (!c || !c->sand.ready) return set_error(LUS_ESTATE, "lus_sand_set_graph: no sand scene");
    SandState& S = c->sand;
    if (on && S.fixed_cap <= 0)
        return set_error(LUS_ESTATE, "lus_sand_set_graph: needs a fixed list capacity (lus_sand_set_list_capacity)");
    LUS_HIP(hipSetDevice(c->device));
    LUS_HIP(hipStreamSynchronize(c->stream));
    sand_drop_graphs(S);
    S.use_graph = on != 0;
    return LUS_OK;
}

int lus_sand_set_variant(LusGpu* c, int variant) {
    if (!c || !c->sand.ready) return set_error(LUS_ESTATE, "lus_sand_set_variant: no sand scene");
    if (variant < 0 || variant > 2) return set_error(LUS_EINVAL, "lus_sand_set_variant: variant %d", variant);
    SandState& S = c->sand;
    LUS_HIP(hipSetDevice(c->device));
    LUS_HIP(hipStreamSynchronize(c->stream));
    sand_drop_graphs(S);  // the constants are kernel arguments of the captured steps
    lus::SandK& k = S.k;
    // the locals of each function (src/Simulate.cpp:159-163, 330-334, 516-520)
    static const float coeff[3][4] = {{0.8f, 0.7f, 0.95f, 0.9f}, {0.8f, 0.7f, 0.8f, 0.7f}, {0.8f, 0.5f, 0.8f, 0.8f}};
    k.variant = variant;
    k.collision = coeff[variant][0];
    k.friction = coeff[variant][1];
    k.mu_s = coeff[variant][2];
    k.mu_k = coeff[variant][3];
    k.bcoef = 0.9f;
    S.p.collision_coeff = k.collision;
    S.p.friction_coeff = k.friction;
    S.p.mu_s = k.mu_s;
    S.p.mu_k = k.mu_k;
    return LUS_OK;
}

int lus_sand_download(LusGpu* c, float* pos, float* pos_star, float* vel, int* attracted, int* perm) {
    if (!c || !c->sand.ready) return set_error(LUS_ESTATE, "lus_sand_download: no sand scene");
    SandState& S = c->sand;
    if (perm && !S.stepped) return set_error(LUS_ESTATE, "lus_sand_download: no step yet");
    LUS_HIP(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    const int n = S.n;
    // vec3 rows are packed on the device and each array is one copy into the
    // caller's memory; the stream is waited for once
    if ((pos || pos_star || vel) && !S.stage) LUS_HIP(S.mem.alloc(&S.stage, 6 * (size_t)n));
    if (pos || pos_star) {
        LUS_HIP(lus::launch_pack3(S.x, nullptr, n, S.stage, s));
        LUS_HIP(hipMemcpyAsync(pos ? pos : pos_star, S.stage, sizeof(float) * 3 * n, hipMemcpyDeviceToHost, s));
    }
    if (vel) {
        LUS_HIP(lus::launch_pack3(S.v, nullptr, n, S.stage + 3 * (size_t)n, s));
        LUS_HIP(hipMemcpyAsync(vel, S.stage + 3 * (size_t)n, sizeof(float) * 3 * n, hipMemcpyDeviceToHost, s));
    }
    if (attracted) LUS_HIP(hipMemcpyAsync(attracted, S.att, sizeof(int) * n, hipMemcpyDeviceToHost, s));
    if (perm) LUS_HIP(hipMemcpyAsync(perm, S.perm, sizeof(int) * n, hipMemcpyDeviceToHost, s));
    LUS_HIP(hipStreamSynchronize(s));
    if (int r = sand_check(c)) return r;
    // positions_star equals positions after a step (src/Simulate.cpp:316-319)
    // and before the first one (both are the uploaded positions)
    if (pos && pos_star) memcpy(pos_star, pos, sizeof(float) * 3 * (size_t)n);
    return LUS_OK;
}

// get_cell_indices (src/neighbors/Utils.hpp:13-21): clamp to
// [cs/2, domain - cs/2], divide by the cell size, truncate
static void sand_cell_indices(const LusSandParams& p, const float* pos, int* out) {
    const float cs = p.grid.cell_size;
    const float half = cs * 0.5f;
    for (int a = 0; a < 3; ++a) {
        const float hi = p.domain[a] - half;
        float q = pos[a] < half ? half : pos[a];
        q = q > hi ? hi : q;
        volatile float d = q / cs;
        out[a] = (int)d;
    }
}

int lus_sand_query_cells(LusGpu* c, const float* min_pos, const float* max_pos, int include_solid, int* count) {
    if (!c || !c->sand.ready) return set_error(LUS_ESTATE, "lus_sand_query_cells: no sand scene");
    if (!min_pos || !max_pos || !count) return set_error(LUS_EINVAL, "lus_sand_query_cells: bad arguments");
    SandState& S = c->sand;
    *count = 0;
    // before the first step the reference's cell lists are empty
    if (!S.stepped) return LUS_OK;
    int lo[3], hi[3];
    sand_cell_indices(S.p, min_pos, lo);
    sand_cell_indices(S.p, max_pos, hi);
    if (hi[0] < lo[0] || hi[1] < lo[1] || hi[2] < lo[2]) return LUS_OK;
    LUS_HIP(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    LUS_HIP(hipMemsetAsync(S.qbuf, 0, sizeof(int), s));
    LUS_HIP(lus::launch_cell_box_count(S.cs, S.ocs, S.solid_ref, lo, hi, S.k.gX, S.k.gZ, S.n,
                                       include_solid ? 1 : 0, S.qbuf, s));
    LUS_HIP(hipMemcpyAsync(count, S.qbuf, sizeof(int), hipMemcpyDeviceToHost, s));
    LUS_HIP(hipStreamSynchronize(s));
    return LUS_OK;
}

int lus_sand_sink_evict(LusGpu* c, const int* cells, int ncells, int* n_sand) {
    if (!c || !c->sand.ready) return set_error(LUS_ESTATE, "lus_sand_sink_evict: no sand scene");
    if (ncells < 0 || (ncells > 0 && !cells)) return set_error(LUS_EINVAL, "lus_sand_sink_evict: bad arguments");
    SandState& S = c->sand;
    for (int k = 0; k < ncells; ++k)
        if (cells[k] < 0 || cells[k] >= S.C)
            return set_error(LUS_EINVAL, "lus_sand_sink_evict: cell %d (entry %d) outside the grid", cells[k], k);
    if (S.stepped && ncells > 0) {
        LUS_HIP(hipSetDevice(c->device));
        hipStream_t s = c->stream;
        Arena tmp;
        int* d_cells;
        int2* d_rng;
        LUS_HIP(tmp.alloc(&d_cells, ncells));
        LUS_HIP(tmp.alloc(&d_rng, ncells));
        LUS_HIP(hipMemcpyAsync(d_cells, cells, sizeof(int) * ncells, hipMemcpyHostToDevice, s));
        LUS_HIP(lus::launch_cell_ranges(S.cs, d_cells, ncells, d_rng, s));
        std::vector<int2> rng((size_t)ncells);
        LUS_HIP(hipMemcpyAsync(rng.data(), d_rng, sizeof(int2) * ncells, hipMemcpyDeviceToHost, s));
        LUS_HIP(hipStreamSynchronize(s));
        // temp_removal: the sand slots of the listed cells in list order (each
        // cell's static solids come first in the reference's lists; their
        // slots lie above the sand range there, so none is taken)
        std::vector<int> rem;
        for (int k = 0; k < ncells; ++k)
            for (int q = rng[k].x; q < rng[k].y; ++q)
                if (q < S.n) rem.push_back(q);
        // the reference's removal, one after another: slot e takes the data of
        // the last sand slot and the range shrinks by one.  Replayed on slot
        // indices; only the final content of each touched slot is moved.
        std::unordered_map<int, int> src;
        auto origin = [&](int q) {
            auto it = src.find(q);
            return it == src.end() ? q : it->second;
        };
        int end = S.n;
        for (int e : rem) {
            if (end == 0) break;
            const int from = origin(end - 1);
            src[e] = from;
            --end;
        }
        std::vector<int2> mv;
        for (const auto& kv : src)
            if (kv.first < end && kv.second != kv.first) mv.push_back(make_int2(kv.first, kv.second));
        const int m = (int)mv.size();
        if (m > 0) {
            int2* d_mv;
            float4 *t0, *t1, *t2;
            int* t3;
            LUS_HIP(tmp.alloc(&d_mv, m));
            LUS_HIP(tmp.alloc(&t0, m));
            LUS_HIP(tmp.alloc(&t1, m));
            LUS_HIP(tmp.alloc(&t2, m));
            LUS_HIP(tmp.alloc(&t3, m));
            LUS_HIP(hipMemcpyAsync(d_mv, mv.data(), sizeof(int2) * m, hipMemcpyHostToDevice, s));
            LUS_HIP(lus::launch_slot_moves(d_mv, m, S.x, S.final_iter, S.v, S.att, t0, t1, t2, t3, s));
            LUS_HIP(hipStreamSynchronize(s));
        }
        S.n = end;
    }
    if (n_sand) *n_sand = S.n;
    return LUS_OK;
}

int lus_sand_player_box(LusGpu* c, const float* player, const float* half_extent, int max_out, int* slots,
                        float* pos, int* count) {
    if (!c || !c->sand.ready) return set_error(LUS_ESTATE, "lus_sand_player_box: no sand scene");
    if (!player || !half_extent || !count || max_out < 0)
        return set_error(LUS_EINVAL, "lus_sand_player_box: bad arguments");
    SandState& S = c->sand;
    *count = 0;
    const int n = S.n;
    if (n == 0 || max_out == 0) return LUS_OK;
    LUS_HIP(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    const int nb = (n + lus::kBlock - 1) / lus::kBlock;
    Arena tmp;
    int *bc, *off, *tiles, *d_slots;
    float4* d_pos;
    LUS_HIP(tmp.alloc(&bc, nb + 1));
    LUS_HIP(tmp.alloc(&off, nb + 1));
    LUS_HIP(tmp.alloc(&tiles, lus::scan_tile_count(nb + 1) + 1));
    LUS_HIP(tmp.alloc(&d_slots, max_out));
    LUS_HIP(tmp.alloc(&d_pos, max_out));
    LUS_HIP(hipMemsetAsync(bc + nb, 0, sizeof(int), s));
    LUS_HIP(lus::launch_box_count(S.x, n, player, half_extent, bc, s));
    LUS_HIP(lus::launch_exclusive_scan(bc, off, nb + 1, tiles, 0, s));
    LUS_HIP(lus::launch_box_write(S.x, n, player, half_extent, off, max_out, d_slots, d_pos, s));
    int total = 0;
    LUS_HIP(hipMemcpyAsync(&total, off + nb, sizeof(int), hipMemcpyDeviceToHost, s));
    LUS_HIP(hipStreamSynchronize(s));
    const int k = std::min(total, max_out);
    if (k > 0 && slots) LUS_HIP(hipMemcpy(slots, d_slots, sizeof(int) * k, hipMemcpyDeviceToHost));
    if (k > 0 && pos) {
        std::vector<float4> h((size_t)k);
        LUS_HIP(hipMemcpy(h.data(), d_pos, sizeof(float4) * k, hipMemcpyDeviceToHost));
        from_f4(h, pos);
    }
    *count = k;
    return LUS_OK;
}

int lus_chunk_from_grid(LusGpu* c, const int* cells, int X, int Y, int Z, const float* position, float cell_size,
                        int subdivision, int mask_out, int num_occupied, const float* cell_colors, float* positions,
                        float* colors, int* generated) {
    if (!c || !cells || !position || !generated || X <= 0 || Y <= 0 || Z <= 0 || subdivision <= 0 ||
        num_occupied < 0)
        return set_error(LUS_EINVAL, "lus_chunk_from_grid: bad arguments");
    const long ncell = (long)X * Y * Z;
    const long s3 = (long)subdivision * subdivision * subdivision;
    const long F = ncell * s3;
    const long cap = (long)num_occupied * s3;
    if (F >= INT_MAX || cap >= INT_MAX || ncell >= INT_MAX)
        return set_error(LUS_EINVAL, "lus_chunk_from_grid: %ld lattice points, at most %d", F, INT_MAX - 1);
    if (cap > 0 && !positions) return set_error(LUS_EINVAL, "lus_chunk_from_grid: no positions buffer");
    const bool want_col = cell_colors && colors;
    LUS_HIP(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    Arena tmp;
    int *d_cells, *flags, *rank, *tiles;
    float4* d_ccol = nullptr;
    LUS_HIP(tmp.alloc(&d_cells, ncell));
    LUS_HIP(tmp.alloc(&flags, F + 1));
    LUS_HIP(tmp.alloc(&rank, F + 1));
    LUS_HIP(tmp.alloc(&tiles, lus::scan_tile_count((int)F + 1) + 1));
    LUS_HIP(hipMemcpyAsync(d_cells, cells, sizeof(int) * ncell, hipMemcpyHostToDevice, s));
    if (want_col) {
        std::vector<float4> hc((size_t)ncell);
        for (long q = 0; q < ncell; ++q)
            hc[q] = make_float4(cell_colors[4 * q], cell_colors[4 * q + 1], cell_colors[4 * q + 2], cell_colors[4 * q + 3]);
        LUS_HIP(tmp.alloc(&d_ccol, ncell));
        LUS_HIP(hipMemcpyAsync(d_ccol, hc.data(), sizeof(float4) * ncell, hipMemcpyHostToDevice, s));
        LUS_HIP(hipStreamSynchronize(s));
    }
    LUS_HIP(hipMemsetAsync(flags + F, 0, sizeof(int), s));
    LUS_HIP(lus::launch_chunk_flags(d_cells, X, Y, Z, subdivision, mask_out ? 1 : 0, F, flags, s));
    LUS_HIP(lus::launch_exclusive_scan(flags, rank, (int)F + 1, tiles, 0, s));
    int total = 0;
    LUS_HIP(hipMemcpyAsync(&total, rank + F, sizeof(int), hipMemcpyDeviceToHost, s));
    LUS_HIP(hipStreamSynchronize(s));
    // the reference sizes the chunk by num_occupied_grid_cells and writes one
    // entry per kept point: more points than that would write past its end
    if (total > cap)
        return set_error(LUS_EINVAL, "lus_chunk_from_grid: %d points for %ld entries (num_occupied too small)", total,
                         cap);
    float4 *d_pos = nullptr, *d_col = nullptr;
    LUS_HIP(tmp.alloc(&d_pos, total));
    if (want_col) LUS_HIP(tmp.alloc(&d_col, total));
    LUS_HIP(lus::launch_chunk_write(d_cells, d_ccol, X, Y, Z, subdivision, mask_out ? 1 : 0, F, position, cell_size,
                                    rank, d_pos, d_col, s));
    std::vector<float4> h((size_t)std::max(total, 1));
    LUS_HIP(hipMemcpyAsync(h.data(), d_pos, sizeof(float4) * total, hipMemcpyDeviceToHost, s));
    LUS_HIP(hipStreamSynchronize(s));
    for (int q = 0; q < total; ++q) {
        positions[3 * (size_t)q] = h[q].x;
        positions[3 * (size_t)q + 1] = h[q].y;
        positions[3 * (size_t)q + 2] = h[q].z;
    }
    if (cap > total) memset(positions + 3 * (size_t)total, 0, sizeof(float) * 3 * (size_t)(cap - total));
    if (want_col) {
        LUS_HIP(hipMemcpy(h.data(), d_col, sizeof(float4) * total, hipMemcpyDeviceToHost));
        memcpy(colors, h.data(), sizeof(float4) * (size_t)total);
        if (cap > total) memset(colors + 4 * (size_t)total, 0, sizeof(float) * 4 * (size_t)(cap - total));
    }
    *generated = total;
    return LUS_OK;
}

int lus_grid_from_voxels(LusGpu* c, const unsigned char* voxels, int sx, int sy, int sz,
                         const unsigned char* palette, int* cells, float* colors, int* occupied) {
    if (!c || !voxels || !palette || !cells || !occupied || sx <= 0 || sy <= 0 || sz <= 0)
        return set_error(LUS_EINVAL, "lus_grid_from_voxels: bad arguments");
    const long N = (long)sx * sy * sz;
    if (N >= INT_MAX) return set_error(LUS_EINVAL, "lus_grid_from_voxels: model too large");
    LUS_HIP(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    Arena tmp;
    unsigned char* d_vox;
    uchar4* d_pal;
    int *d_cells, *d_occ;
    float4* d_col;
    LUS_HIP(tmp.alloc(&d_vox, N));
    LUS_HIP(tmp.alloc(&d_pal, 256));
    LUS_HIP(tmp.alloc(&d_cells, N));
    LUS_HIP(tmp.alloc(&d_col, N));
    LUS_HIP(tmp.alloc(&d_occ, 1));
    LUS_HIP(hipMemcpyAsync(d_vox, voxels, N, hipMemcpyHostToDevice, s));
    LUS_HIP(hipMemcpyAsync(d_pal, palette, 256 * 4, hipMemcpyHostToDevice, s));
    LUS_HIP(hipMemsetAsync(d_occ, 0, sizeof(int), s));
    LUS_HIP(lus::launch_grid_from_voxels(d_vox, sx, sy, sz, d_pal, d_cells, d_col, d_occ, s));
    LUS_HIP(hipMemcpyAsync(cells, d_cells, sizeof(int) * N, hipMemcpyDeviceToHost, s));
    if (colors) LUS_HIP(hipMemcpyAsync(colors, d_col, sizeof(float4) * N, hipMemcpyDeviceToHost, s));
    LUS_HIP(hipMemcpyAsync(occupied, d_occ, sizeof(int), hipMemcpyDeviceToHost, s));
    LUS_HIP(hipStreamSynchronize(s));
    return LUS_OK;
}

int lus_sand_neighbors(LusGpu* c, long* off, int* idx, long cap, long* total) {
    if (!c || !c->sand.ready || !c->sand.stepped || !off || !total)
        return set_error(LUS_ESTATE, "lus_sand_neighbors: no step yet");
    SandState& S = c->sand;
    LUS_HIP(hipSetDevice(c->device));
    LUS_HIP(hipStreamSynchronize(c->stream));
    if (int r = sand_check(c)) return r;
    const int n = S.n;
    std::vector<int> o((size_t)n + 1);
    LUS_HIP(hipMemcpy(o.data(), S.nbr_off, sizeof(int) * (n + 1), hipMemcpyDeviceToHost));
    for (int i = 0; i <= n; ++i) off[i] = o[i];
    *total = o[n];
    if (*total > cap || (*total > 0 && !idx))
        return set_error(LUS_EINVAL, "lus_sand_neighbors: capacity %ld < %ld", cap, *total);
    if (*total == 0) return LUS_OK;
    LUS_HIP(hipMemcpy(idx, S.nbr_idx, sizeof(int) * (size_t)*total, hipMemcpyDeviceToHost));
    std::vector<int> ref((size_t)std::max(S.ns, 1));
    if (S.ns) LUS_HIP(hipMemcpy(ref.data(), S.solid_ref, sizeof(int) * S.ns, hipMemcpyDeviceToHost));
    for (long t = 0; t < *total; ++t)
        if (idx[t] >= n) idx[t] = ref[idx[t] - n];
    return LUS_OK;
}

}  // extern "C"
