// lus_capi.cpp -- host side of the C ABI declared in include/lustrine_gpu.h:
// argument checks, device buffers, the per-step launch sequences and timing.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <vector>

#include "lus_device.hpp"
#include "lus_internal.hpp"
#include "lustrine_gpu.h"

namespace {

int g_err_code = 0;
char g_err_msg[512] = "";

int set_error(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err_msg, sizeof(g_err_msg), fmt, ap);
    va_end(ap);
    g_err_code = code;
    fprintf(stderr, "lustrine_gpu: %s\n", g_err_msg);
    return code;
}

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
    float4 *x = nullptr, *x2 = nullptr, *v = nullptr, *v2 = nullptr, *xa = nullptr, *xl = nullptr,
           *xp = nullptr;
    int *id = nullptr, *id2 = nullptr, *key = nullptr, *slot = nullptr, *tmp = nullptr,
        *key2 = nullptr, *counts = nullptr, *cs = nullptr, *tiles = nullptr, *maxc = nullptr;
    unsigned long long* masks = nullptr;  // 9 neighbour-row masks per particle
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
    bool stepped = false;
};

struct LusGpu {
    int device = 0;
    hipStream_t stream = nullptr;
    PbfState pbf;
    SandState sand;
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
    k.n_scorr = p.scorr_n;
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

// Enqueues one PBF substep.  ev (nullable): LUS_PBF_NPHASES-indexed event log,
// ev_phase[k] receives the phase of event k; returns the number of events.
int pbf_enqueue(LusGpu* c, float dt, std::vector<hipEvent_t>* ev, std::vector<int>* ev_phase) {
    PbfState& P = c->pbf;
    hipStream_t s = c->stream;
    auto mark = [&](int phase) -> hipError_t {
        if (!ev) return hipSuccess;
        hipEvent_t e;
        hipError_t r = hipEventCreate(&e);
        if (r != hipSuccess) return r;
        ev->push_back(e);
        ev_phase->push_back(phase);
        return hipEventRecord(e, s);
    };
    // per-step constants: the scene-id buffers of a batch ping-pong like id
    lus::PbfK k = P.k;
    if (P.scenes) {
        k.scenes = P.scenes;
        k.sid = P.sid;
        k.sid_out = P.sid2;
    }
    if (ev) LUS_HIP(mark(-1));
    LUS_HIP(lus::launch_pbf_key(P.x, P.v, P.n, k, dt, P.key, P.slot, P.counts, s));
    LUS_HIP(mark(0));
    LUS_HIP(lus::launch_exclusive_scan(P.counts, P.cs, P.C + 1, P.tiles, 1, s));
    LUS_HIP(mark(1));
    LUS_HIP(lus::launch_scatter(P.key, P.slot, P.cs, P.n, P.tmp, s));
    LUS_HIP(mark(2));
    LUS_HIP(lus::launch_pbf_permute(P.tmp, P.key, P.cs, P.x, P.v, P.id, P.n, k, dt, P.x2, P.xp,
                                    P.id2, P.key2, s));
    LUS_HIP(mark(3));
    // iterate: xp (predicted) -> lambda -> xl -> delta-p -> xa -> lambda -> xl ...
    const int K = P.p.iterations;
    for (int it = 0; it < K; ++it) {
        const float4* cur = it == 0 ? P.xp : P.xa;
        LUS_HIP(lus::launch_pbf_lambda(cur, P.key2, P.cs, P.n, k, P.xl, P.maxc + (size_t)it * P.nblk,
                                       P.masks, P.xp, it == 0 ? 1 : 0, s));
        LUS_HIP(mark(it == 0 ? 7 : 4));
        const bool last = (it == K - 1);
        LUS_HIP(lus::launch_pbf_deltap(P.xl, P.key2, P.cs, P.n, k, P.xa, P.x2,
                                       last ? P.v2 : nullptr, dt, P.masks, P.xp, s));
        LUS_HIP(mark(5));
    }
    const float4* fin = K > 0 ? P.xa : P.xp;
    if (K == 0) {
        // neighbour sets for XSPH (the lambda values are not used)
        LUS_HIP(lus::launch_pbf_lambda(P.xp, P.key2, P.cs, P.n, k, P.xl, P.maxc, P.masks, P.xp, 1, s));
        LUS_HIP(lus::launch_pbf_velocity(P.xp, P.x2, P.n, dt, P.v2, s));
    }
    LUS_HIP(lus::launch_pbf_xsph(fin, P.v2, P.key2, P.cs, P.n, k, P.v, P.x2, P.masks, P.xp, s));
    LUS_HIP(mark(6));
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
    c->pbf.mem.release();
    c->sand.mem.release();
    c->sand.nbr_mem.release();
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
    return LUS_OK;
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

}  // extern "C"

namespace {
int pbf_setup(LusGpu* c, const LusPbfParams* p, int n, const float* pos, const float* vel,
              const std::vector<lus::SceneBox>* scenes, const std::vector<int>* sid) {
    LUS_HIP(hipSetDevice(c->device));
    PbfState& P = c->pbf;
    LUS_HIP(hipStreamSynchronize(c->stream));
    pbf_drop_graph(P);
    P.mem.release();
    P.ready = false;
    P.stepped = false;
    P.parity = 0;
    P.scenes = nullptr;
    P.sid = P.sid2 = nullptr;
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
    LUS_HIP(m.alloc(&P.masks, (size_t)9 * n));
    LUS_HIP(m.alloc(&P.id, n));
    LUS_HIP(m.alloc(&P.id2, n));
    LUS_HIP(m.alloc(&P.key, n));
    LUS_HIP(m.alloc(&P.slot, n));
    LUS_HIP(m.alloc(&P.tmp, n));
    LUS_HIP(m.alloc(&P.key2, n));
    LUS_HIP(m.alloc(&P.counts, P.C + 1));
    LUS_HIP(m.alloc(&P.cs, P.C + 1));
    LUS_HIP(m.alloc(&P.tiles, lus::scan_tile_count(P.C + 1) + 1));
    P.nblk = lus::grid_for(n);
    LUS_HIP(m.alloc(&P.maxc, (size_t)P.nmaxc * P.nblk));
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
    return pbf_run(c, dt, steps);
}

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

int lus_pbf_download(LusGpu* c, int order, float* pos, float* vel, int* ids) {
    if (!c || !c->pbf.ready) return set_error(LUS_ESTATE, "lus_pbf_download: no PBF scene");
    PbfState& P = c->pbf;
    LUS_HIP(hipSetDevice(c->device));
    LUS_HIP(hipStreamSynchronize(c->stream));
    const int n = P.n;
    std::vector<float4> hx((size_t)n), hv((size_t)n);
    std::vector<int> hid((size_t)n);
    LUS_HIP(hipMemcpy(hx.data(), P.x, sizeof(float4) * n, hipMemcpyDeviceToHost));
    LUS_HIP(hipMemcpy(hv.data(), P.v, sizeof(float4) * n, hipMemcpyDeviceToHost));
    LUS_HIP(hipMemcpy(hid.data(), P.id, sizeof(int) * n, hipMemcpyDeviceToHost));
    if (order == 1) {
        std::vector<float4> ox((size_t)n), ov((size_t)n);
        for (int r = 0; r < n; ++r) {
            ox[hid[r]] = hx[r];
            ov[hid[r]] = hv[r];
        }
        hx.swap(ox);
        hv.swap(ov);
        for (int r = 0; r < n; ++r) hid[r] = r;
    } else if (order != 0) {
        return set_error(LUS_EINVAL, "lus_pbf_download: order must be 0 or 1");
    }
    if (pos) from_f4(hx, pos);
    if (vel) from_f4(hv, vel);
    if (ids) memcpy(ids, hid.data(), sizeof(int) * n);
    return LUS_OK;
}

int lus_pbf_last_max_constraint(LusGpu* c, float* out) {
    if (!c || !c->pbf.ready || !out) return set_error(LUS_ESTATE, "lus_pbf_last_max_constraint: no scene");
    PbfState& P = c->pbf;
    LUS_HIP(hipStreamSynchronize(c->stream));
    int last = std::max(P.p.iterations - 1, 0);
    std::vector<int> blk((size_t)P.nblk);
    LUS_HIP(hipMemcpy(blk.data(), P.maxc + (size_t)last * P.nblk, sizeof(int) * P.nblk,
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
        for (int s = 0; s < steps; ++s) {
            std::vector<hipEvent_t> ev;
            std::vector<int> ph;
            r = pbf_enqueue(c, dt, &ev, &ph);
            if (r) return r;
            LUS_HIP(hipEventSynchronize(ev.back()));
            for (size_t k = 1; k < ev.size(); ++k) {
                float t = 0.0f;
                LUS_HIP(hipEventElapsedTime(&t, ev[k - 1], ev[k]));
                phase_ms[ph[k]] += t;
                *ms += t;
                if (phase_launches) phase_launches[ph[k]] += 1;
            }
            for (hipEvent_t e : ev) (void)hipEventDestroy(e);
        }
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
    for (int kind = 0; kind < 4; ++kind) {
        LUS_HIP(hipStreamSynchronize(s));
        LUS_HIP(hipEventRecord(e0, s));
        for (int r = 0; r < reps; ++r) {
            // xl, xa and x2 are per-step scratch between steps; x, v, id hold the state
            if (kind == 0)
                LUS_HIP(lus::launch_pbf_lambda(P.xa, P.key2, P.cs, P.n, P.k, P.xl, P.maxc, P.masks, P.xp,
                                               0, s));
            else if (kind == 1)
                LUS_HIP(lus::launch_pbf_deltap(P.xl, P.key2, P.cs, P.n, P.k, P.xa, nullptr, nullptr,
                                               1.0f, P.masks, P.xp, s));
            else if (kind == 2)
                LUS_HIP(lus::launch_pbf_xsph(P.xa, P.v2, P.key2, P.cs, P.n, P.k, P.xl, P.x2, P.masks,
                                             P.xp, s));
            else
                LUS_HIP(lus::launch_pbf_lambda(P.xp, P.key2, P.cs, P.n, P.k, P.xl, P.maxc, P.masks, P.xp,
                                               1, s));
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
    S.mem.release();
    S.nbr_mem.release();
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
    std::vector<float4> hx = to_f4(pos, n), hv = to_f4(vel, n);
    std::vector<int> ha((size_t)n, 0);
    if (attracted) memcpy(ha.data(), attracted, sizeof(int) * n);
    LUS_HIP(hipMemcpy(S.x, hx.data(), sizeof(float4) * n, hipMemcpyHostToDevice));
    LUS_HIP(hipMemcpy(S.v, hv.data(), sizeof(float4) * n, hipMemcpyHostToDevice));
    LUS_HIP(hipMemcpy(S.att, ha.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    LUS_HIP(hipMemset(S.counts, 0, sizeof(int) * (S.C + 1)));
    LUS_HIP(hipMemset(S.nbr_cnt, 0, sizeof(int) * (n + 1)));
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

int lus_sand_step(LusGpu* c, float dt, int attract_flag, int blow_flag, const float* player) {
    if (!c || !c->sand.ready) return set_error(LUS_ESTATE, "lus_sand_step: no sand scene");
    if (!(dt != 0.0f) || ((attract_flag || blow_flag) && !player))
        return set_error(LUS_EINVAL, "lus_sand_step: bad arguments");
    LUS_HIP(hipSetDevice(c->device));
    SandState& S = c->sand;
    hipStream_t s = c->stream;
    const int n = S.n;
    lus::SandK k = S.k;
    k.attract = attract_flag ? 1 : 0;
    k.blow = blow_flag ? 1 : 0;
    k.prev_attract = S.prev_attract;
    for (int a = 0; a < 3; ++a) k.player[a] = player ? player[a] : 0.0f;
    LUS_HIP(lus::launch_sand_predict_key(S.x, S.v, S.att, n, k, dt, S.xs, S.key, S.slot, S.counts, s));
    LUS_HIP(lus::launch_exclusive_scan(S.counts, S.cs, S.C + 1, S.tiles, 1, s));
    LUS_HIP(lus::launch_scatter(S.key, S.slot, S.cs, n, S.tmp, s));
    LUS_HIP(lus::launch_sand_permute(S.tmp, S.key, S.cs, S.x, S.xs, S.v, S.att, n, S.x2, S.A, S.v2,
                                     S.att2, S.perm, S.key2, s));
    // neighbor lists: count, scan, fill
    LUS_HIP(lus::launch_sand_nbr(S.A, S.key2, S.cs, n, S.solid, S.ocs, k, nullptr, S.nbr_cnt,
                                 S.nbr_cnt1, nullptr, 0, s));
    LUS_HIP(lus::launch_exclusive_scan(S.nbr_cnt, S.nbr_off, n + 1, S.ntiles, 0, s));
    int total = 0;
    LUS_HIP(hipMemcpyAsync(&total, S.nbr_off + n, sizeof(int), hipMemcpyDeviceToHost, s));
    LUS_HIP(hipStreamSynchronize(s));
    if (total < 0) return set_error(LUS_EHIP, "lus_sand_step: neighbor count overflow");
    if ((long)total > S.nbr_cap || !S.nbr_idx) {
        S.nbr_mem.release();
        long cap = std::max<long>((long)total + total / 4, 64);
        LUS_HIP(S.nbr_mem.alloc(&S.nbr_idx, (size_t)cap));
        S.nbr_cap = cap;
    }
    LUS_HIP(lus::launch_sand_nbr(S.A, S.key2, S.cs, n, S.solid, S.ocs, k, S.nbr_off, S.nbr_cnt,
                                 S.nbr_cnt1, S.nbr_idx, 1, s));
    float4* A = S.A;
    float4* B = S.B;
    for (int it = 0; it < S.p.substeps; ++it) {
        LUS_HIP(lus::launch_sand_substep(A, S.x2, S.solid, S.nbr_off, S.nbr_idx, n, k, B, s));
        std::swap(A, B);
    }
    LUS_HIP(lus::launch_sand_vupdate(A, S.x2, S.v2, n, dt, s));
    S.A = A;
    S.B = B;
    S.final_iter = A;
    std::swap(S.x, S.x2);
    std::swap(S.v, S.v2);
    std::swap(S.att, S.att2);
    S.prev_attract = k.attract;
    S.stepped = true;
    return LUS_OK;
}

int lus_sand_download(LusGpu* c, float* pos, float* pos_star, float* vel, int* attracted, int* perm) {
    if (!c || !c->sand.ready) return set_error(LUS_ESTATE, "lus_sand_download: no sand scene");
    SandState& S = c->sand;
    LUS_HIP(hipSetDevice(c->device));
    LUS_HIP(hipStreamSynchronize(c->stream));
    const int n = S.n;
    std::vector<float4> h((size_t)n);
    if (pos) {
        LUS_HIP(hipMemcpy(h.data(), S.x, sizeof(float4) * n, hipMemcpyDeviceToHost));
        from_f4(h, pos);
    }
    if (pos_star) {
        LUS_HIP(hipMemcpy(h.data(), S.stepped ? S.final_iter : S.x, sizeof(float4) * n,
                          hipMemcpyDeviceToHost));
        from_f4(h, pos_star);
    }
    if (vel) {
        LUS_HIP(hipMemcpy(h.data(), S.v, sizeof(float4) * n, hipMemcpyDeviceToHost));
        from_f4(h, vel);
    }
    if (attracted) LUS_HIP(hipMemcpy(attracted, S.att, sizeof(int) * n, hipMemcpyDeviceToHost));
    if (perm) {
        if (!S.stepped) return set_error(LUS_ESTATE, "lus_sand_download: no step yet");
        LUS_HIP(hipMemcpy(perm, S.perm, sizeof(int) * n, hipMemcpyDeviceToHost));
    }
    return LUS_OK;
}

int lus_sand_neighbors(LusGpu* c, long* off, int* idx, long cap, long* total) {
    if (!c || !c->sand.ready || !c->sand.stepped || !off || !total)
        return set_error(LUS_ESTATE, "lus_sand_neighbors: no step yet");
    SandState& S = c->sand;
    LUS_HIP(hipSetDevice(c->device));
    LUS_HIP(hipStreamSynchronize(c->stream));
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
