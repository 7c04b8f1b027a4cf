// lus_fluid.hip -- the reference's simulate_fluid (src/Simulate.cpp:27-115) on
// the GPU, with the reference's semantics kept exactly:
//
//   * dt clamped to [0.001, 0.01]; v += g * mass * dt; x* = x + v dt (:31, :48-51)
//   * neighbour lists of find_neighbors_uniform_grid (Neighbors.cpp:366-450):
//     per particle the 27 cells in (dy, dx, dz) order, each cell's moving
//     particles in ascending slot then its solids in ascending slot, self
//     included; built from a stable counting sort of the predicted positions
//   * lambda per particle with the cubic kernel (Kernels.cpp:5-41), self
//     counted twice in the density (:58-88) -- independent per particle
//   * the pressure sweep (:90-113) is sequential in the reference: particle i
//     reads x* of the particles before it as already corrected (in place) and
//     of those after it as predicted, and uses lambdas[t] for the t-th entry
//     of its list.  Here it runs as a dependency-level schedule: particle i
//     waits for every neighbour j < i; the moving particles whose waits are
//     all met form one level and run in one launch, and a launch boundary
//     separates the levels.  Two particles of one level are never neighbours
//     (the later one would wait for the earlier), so a particle reads every
//     earlier neighbour already corrected and every later one not yet
//     touched: exactly the reference's sequential sweep, in parallel.
//
// Arithmetic follows the reference's operation order with no contraction and
// IEEE divide and square root (the library's build flags).  One call differs
// from the reference's binary: std::pow(float, float) in s_coor is glibc's
// powf; here the power is taken in double and rounded once (the correctly
// rounded value), which glibc's powf returns except for about 7 in 10,000
// arguments, where it is one ulp away.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <vector>

#include "lus_device.hpp"
#include "lus_internal.hpp"
#include "lus_launch.hpp"
#include "lustrine_gpu.h"

namespace lus {

struct FluidK {
    float r, h, h2, f, ck, cl, mass, rho0, eps, scorr_k, scorr_n, W0, Wdq;
    float gm[3];          // gravity * mass, per component (float products)
    float hi[3];          // (int)domain - r, per axis: resolve_collision's max
    float cs;
    int gX, gY, gZ, gXZ;
    int s0, n;            // first moving slot, moving count
    int nl;               // entries of the lambdas array
    int ipow;             // s_corr_n as an exact small integer, or -1
};

struct FluidState {
    LusFluidParams p{};
    FluidK k{};
    int total = 0, s0 = 0, s1 = 0, t0 = 0, t1 = 0, C = 0, ns = 0;
    bool stepped = false;
    int last_levels = 0;
    long last_entries = 0;
    std::vector<void*> ptrs;
    float4 *x = nullptr, *xs = nullptr, *v = nullptr, *sorted = nullptr, *osorted = nullptr;
    float* lam = nullptr;
    int *key = nullptr, *slot = nullptr, *counts = nullptr, *cs = nullptr, *tmp = nullptr,
        *perm = nullptr, *ks = nullptr, *tiles = nullptr, *inv = nullptr;
    int *ocs = nullptr, *oref = nullptr;
    int *cnt = nullptr, *off = nullptr, *ltiles = nullptr, *idx = nullptr;
    long cap = 0;
    void* idx_mem = nullptr;
    int *indeg = nullptr, *q0 = nullptr, *q1 = nullptr, *ctl = nullptr;  // ctl: see kCtl*
    int* host_ctl = nullptr;  // pinned mirror of ctl
};

// ctl layout: three frontier counters, processed count, deepest level, error bits
constexpr int kCtlQ = 0;         // ctl[0..2]: frontier counters of levels L % 3
constexpr int kCtlDone = 3;
constexpr int kCtlLevels = 4;
constexpr int kCtlErr = 5;
constexpr int kCtlInts = 8;
constexpr int kErrKey = 1;       // a predicted position lies outside the grid
constexpr int kErrLambda = 2;    // a list is longer than the lambdas array

void fluid_free(FluidState* F) {
    if (!F) return;
    for (void* q : F->ptrs) (void)hipFree(q);
    if (F->idx_mem) (void)hipFree(F->idx_mem);
    if (F->host_ctl) (void)hipHostFree(F->host_ctl);
    delete F;
}

// ------------------------------------------------------------------ kernels
// cubic_kernel (Kernels.cpp:5-19); with kernelFactor 1/2 and neighbours within
// h, q <= 1/2 and only the first branch is taken
__device__ __forceinline__ float fl_cubic(const FluidK& k, float r) {
    const float q = (r * k.f) / k.h;
    float result = 0.0f;
    if ((double)q <= 1.0) {
        if ((double)q <= 0.5) {
            const float q2 = q * q;
            const float q3 = q2 * q;
            const float a = 6.0f * q3;
            const float b = 6.0f * q2;
            result = k.ck * ((a - b) + 1.0f);
        } else {
            const float t = 1.0f - q;
            const double t3 = ((double)t * (double)t) * (double)t;
            result = k.ck * (2.0f * (float)t3);
        }
    }
    return result;
}

// cubic_kernel_grad (Kernels.cpp:26-41)
__device__ __forceinline__ void fl_cubic_grad(const FluidK& k, float dx, float dy, float dz, float* gx,
                                              float* gy, float* gz) {
    const float rl = sqrtf(dot3(dx, dy, dz, dx, dy, dz)) * k.f;
    const float q = rl / k.h;
    float ox = 0.0f, oy = 0.0f, oz = 0.0f;
    if ((double)rl > 1.0e-5 && (double)q <= 1.0) {
        const float c = 1.0f / (rl * k.h);
        const float qx = c * dx, qy = c * dy, qz = c * dz;
        float sc;
        if ((double)q <= 0.5) {
            sc = (k.cl * q) * (3.0f * q - 2.0f);
        } else {
            const float t = 1.0f - q;
            sc = k.cl * ((-t) * t);
        }
        ox = sc * qx;
        oy = sc * qy;
        oz = sc * qz;
    }
    *gx = ox;
    *gy = oy;
    *gz = oz;
}

// std::pow(x, n) of s_coor (Simulate.cpp:7-9), rounded once from double
__device__ __forceinline__ float fl_pow(const FluidK& k, float x) {
    if (k.ipow >= 0) {
        double acc = 1.0, b = (double)x;
        int e = k.ipow;
        while (e) {
            if (e & 1) acc *= b;
            b *= b;
            e >>= 1;
        }
        return (float)acc;
    }
    return (float)pow((double)x, (double)k.scorr_n);
}

// resolve_collision (Simulate.cpp:11-24)
__device__ __forceinline__ float fl_resolve(float value, float lo, float hi) {
    if (value <= lo) return 0.01f;
    if (value > hi) return hi - 0.01f;
    return value;
}

// integration (:48-51) and the cell key of each moving particle
__global__ __launch_bounds__(kBlock) void k_fl_predict(float4* __restrict__ x, float4* __restrict__ v,
                                                        float4* __restrict__ xs, FluidK k, float dt,
                                                        int* __restrict__ key, int C, int* __restrict__ ctl) {
    const int t = xcd_index();
    if (t >= k.n) return;
    const int i = k.s0 + t;
    float4 vi = v[i];
    vi.x = vi.x + k.gm[0] * dt;
    vi.y = vi.y + k.gm[1] * dt;
    vi.z = vi.z + k.gm[2] * dt;
    const float4 p = x[i];
    float4 q;
    q.x = p.x + vi.x * dt;
    q.y = p.y + vi.y * dt;
    q.z = p.z + vi.z * dt;
    q.w = 0.0f;
    v[i] = vi;
    xs[i] = q;
    // get_cell_id (Utils.hpp:24-33) has no clamp: a cell index outside
    // [0, C) indexes outside the reference's cell vector; here it is reported
    // (an index inside [0, C) is used as the reference uses it, even when a
    // coordinate lies past its axis)
    const float px = q.x / k.cs, py = q.y / k.cs, pz = q.z / k.cs;
    bool bad = !(fabsf(px) < 1.0e9f && fabsf(py) < 1.0e9f && fabsf(pz) < 1.0e9f);
    int c = 0;
    if (!bad) {
        const long long cl = (long long)(int)py * k.gXZ + (long long)(int)px * k.gZ + (long long)(int)pz;
        bad = cl < 0 || cl >= C;
        c = (int)cl;
    }
    if (bad) {
        atomicOr(&ctl[kCtlErr], kErrKey);
        c = 0;
    }
    key[t] = c;
}

__global__ __launch_bounds__(kBlock) void k_fl_gather(const float4* __restrict__ xs, const int* __restrict__ perm,
                                                       int s0, int n, float4* __restrict__ sorted,
                                                       int* __restrict__ inv) {
    const int p = xcd_index();
    if (p >= n) return;
    const int t = perm[p];
    sorted[p] = xs[s0 + t];
    inv[t] = p;
}

// find_neighbors_uniform_grid: counts (fill == 0) or entries (fill == 1) of
// moving particle t's list; entries are reference slots
__global__ __launch_bounds__(kBlock) void k_fl_lists(const float4* __restrict__ sorted,
                                                      const int* __restrict__ skey,
                                                      const int* __restrict__ scs,
                                                      const int* __restrict__ sperm,
                                                      const int* __restrict__ inv,
                                                      const float4* __restrict__ osorted,
                                                      const int* __restrict__ ocs,
                                                      const int* __restrict__ oref, FluidK k,
                                                      const int* __restrict__ off, int* __restrict__ cnt,
                                                      int* __restrict__ idx, int fill) {
    const int t = xcd_index();
    if (t >= k.n) return;
    const int p = inv[t];
    const float4 me = sorted[p];
    const int c = skey[p];
    const int yy = c / k.gXZ;
    const int rm = c - yy * k.gXZ;
    const int xx = rm / k.gZ;
    const int zz = rm - xx * k.gZ;
    int o = fill ? off[t] : 0;
    int count = 0;
    for (int y = -1; y <= 1; ++y) {
        if (yy + y < 0 || yy + y >= k.gY) continue;
        for (int x = -1; x <= 1; ++x) {
            if (xx + x < 0 || xx + x >= k.gX) continue;
            for (int z = -1; z <= 1; ++z) {
                if (zz + z < 0 || zz + z >= k.gZ) continue;
                const int nc = (yy + y) * k.gXZ + (xx + x) * k.gZ + (zz + z);
                const int a = scs[nc], b = scs[nc + 1];
                for (int j = a; j < b; ++j) {
                    const float4 q = sorted[j];
                    const float dx = me.x - q.x, dy = me.y - q.y, dz = me.z - q.z;
                    if (dot3(dx, dy, dz, dx, dy, dz) <= k.h2) {
                        if (fill) idx[o + count] = k.s0 + sperm[j];
                        ++count;
                    }
                }
                if (osorted) {
                    const int oa = ocs[nc], ob = ocs[nc + 1];
                    for (int j = oa; j < ob; ++j) {
                        const float4 q = osorted[j];
                        const float dx = me.x - q.x, dy = me.y - q.y, dz = me.z - q.z;
                        if (dot3(dx, dy, dz, dx, dy, dz) <= k.h2) {
                            if (fill) idx[o + count] = oref[j];
                            ++count;
                        }
                    }
                }
            }
        }
    }
    if (!fill) cnt[t] = count;
}

// lambda of each moving particle (:58-88); the number of its neighbours that
// come before it in the sweep; the particles that wait for none start level 0
__global__ __launch_bounds__(kBlock) void k_fl_lambda(const float4* __restrict__ xs, const int* __restrict__ off,
                                                       const int* __restrict__ idx, FluidK k,
                                                       float* __restrict__ lam, int* __restrict__ indeg,
                                                       int* __restrict__ q0, int* __restrict__ ctl) {
    const int t = xcd_index();
    if (t >= k.n) return;
    const int i = k.s0 + t;
    const int a = off[t], b = off[t + 1];
    const float4 pi = xs[i];
    float density = 0.0f;
    int before = 0;
    for (int e = a; e < b; ++e) {
        const int j = idx[e];
        const float4 pj = xs[j];
        const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
        const float len = sqrtf(dot3(dx, dy, dz, dx, dy, dz));
        density = density + k.mass * fl_cubic(k, len);
        before += (j >= k.s0 && j < i) ? 1 : 0;
    }
    density = density + k.mass * k.W0;
    const float constraint = (float)((double)(density / k.rho0) - 1.0);
    float gsum = 0.0f, gcx = 0.0f, gcy = 0.0f, gcz = 0.0f;
    const float m = -(k.mass / k.rho0);
    for (int e = a; e < b; ++e) {
        const int j = idx[e];
        const float4 pj = xs[j];
        float gx, gy, gz;
        fl_cubic_grad(k, pi.x - pj.x, pi.y - pj.y, pi.z - pj.z, &gx, &gy, &gz);
        const float nx = m * gx, ny = m * gy, nz = m * gz;
        gsum = gsum + dot3(nx, ny, nz, nx, ny, nz);
        gcx = gcx - nx;
        gcy = gcy - ny;
        gcz = gcz - nz;
    }
    gsum = gsum + dot3(gcx, gcy, gcz, gcx, gcy, gcz);
    float l = 0.0f;
    if ((double)gsum > 0.0) l = -constraint / (gsum + k.eps);
    lam[i] = l;
    if (b - a > k.nl) atomicOr(&ctl[kCtlErr], kErrLambda);
    indeg[t] = before;
    if (before == 0) q0[atomicAdd(&ctl[kCtlQ + 0], 1)] = t;
}

// One dependency level of the pressure sweep (:90-113).  in / in_cnt: the
// level's particles; out / out_cnt: the next level's, appended here as their
// last wait is met; the counter two levels ahead is cleared for reuse.
__global__ __launch_bounds__(kBlock) void k_fl_sweep(float4* __restrict__ x, float4* __restrict__ xs,
                                                      float4* __restrict__ v, const float* __restrict__ lam,
                                                      const int* __restrict__ off, const int* __restrict__ idx,
                                                      FluidK k, float dt, int level, const int* __restrict__ in,
                                                      int* __restrict__ out, int* __restrict__ indeg,
                                                      int* __restrict__ ctl) {
    const int qi = level % 3;
    const int cnt = ctl[kCtlQ + qi];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ctl[kCtlQ + (level + 2) % 3] = 0;
        if (cnt > 0) atomicMax(&ctl[kCtlLevels], level + 1);
    }
    int* out_cnt = &ctl[kCtlQ + (level + 1) % 3];
    int processed = 0;
    // A level holds about n / depth particles, one per thread: the time of a
    // level is the latency of one particle's walk, so each walk requests its
    // list entries, their positions and lambdas[t] eight at a time before
    // using them (the terms are still added in list order).  The later
    // neighbours waiting for this particle are released in the same walk:
    // they run in a later launch, which sees this launch's writes.
    constexpr int kW = 8;
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < cnt; q += gridDim.x * blockDim.x) {
        const int t = in[q];
        const int i = k.s0 + t;
        const int a = off[t], b = off[t + 1];
        const float4 pi = xs[i];
        const float li = lam[i];
        float fx = 0.0f, fy = 0.0f, fz = 0.0f;
        for (int e0 = a; e0 < b; e0 += kW) {
            int j[kW];
#pragma unroll
            for (int u = 0; u < kW; ++u) j[u] = (e0 + u < b) ? idx[e0 + u] : i;
            float4 pj[kW];
            float lt[kW];
            int old[kW];
#pragma unroll
            for (int u = 0; u < kW; ++u) {
                pj[u] = xs[j[u]];
                // lambdas[j] with j the list position (:97); a list longer
                // than the lambdas array is reported after the step
                const int tp = e0 + u - a;
                lt[u] = (tp < k.nl) ? lam[tp] : 0.0f;
            }
#pragma unroll
            for (int u = 0; u < kW; ++u)
                old[u] = (j[u] > i && j[u] < k.s0 + k.n) ? atomicSub(&indeg[j[u] - k.s0], 1) : 0;
#pragma unroll
            for (int u = 0; u < kW; ++u) {
                if (e0 + u >= b) break;
                const float dx = pi.x - pj[u].x, dy = pi.y - pj[u].y, dz = pi.z - pj[u].z;
                const float len = sqrtf(dot3(dx, dy, dz, dx, dy, dz));
                const float sc = -k.scorr_k * fl_pow(k, fl_cubic(k, len) / k.Wdq);
                const float coef = (li + lt[u]) + sc;
                float gx, gy, gz;
                fl_cubic_grad(k, dx, dy, dz, &gx, &gy, &gz);
                fx = fx + coef * gx;
                fy = fy + coef * gy;
                fz = fz + coef * gz;
            }
#pragma unroll
            for (int u = 0; u < kW; ++u)
                if (j[u] > i && j[u] < k.s0 + k.n && old[u] == 1) out[atomicAdd(out_cnt, 1)] = j[u] - k.s0;
        }
        fx = fx / k.rho0;
        fy = fy / k.rho0;
        fz = fz / k.rho0;
        float4 o;
        o.x = fl_resolve(pi.x + fx, k.r, k.hi[0]);
        o.y = fl_resolve(pi.y + fy, k.r, k.hi[1]);
        o.z = fl_resolve(pi.z + fz, k.r, k.hi[2]);
        o.w = 0.0f;
        xs[i] = o;
        const float4 p0 = x[i];
        float4 vi;
        vi.x = (o.x - p0.x) / dt;
        vi.y = (o.y - p0.y) / dt;
        vi.z = (o.z - p0.z) / dt;
        vi.w = 0.0f;
        v[i] = vi;
        x[i] = o;
        ++processed;
    }
    if (processed) atomicAdd(&ctl[kCtlDone], processed);
}

}  // namespace lus

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
namespace {

#define FL_HIP(call)                                                                                \
    do {                                                                                            \
        hipError_t e_ = (call);                                                                     \
        if (e_ != hipSuccess)                                                                       \
            return lus::report_error(LUS_EHIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                                     __FILE__, __LINE__);                                           \
    } while (0)

template <class T>
hipError_t fl_alloc(lus::FluidState* F, T** out, size_t count) {
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T));
    if (e == hipSuccess) F->ptrs.push_back(q);
    *out = static_cast<T*>(q);
    return e;
}

std::vector<float4> aos_to_f4(const float* xyz, int first, int count) {
    std::vector<float4> out((size_t)std::max(count, 0));
    for (int i = 0; i < count; ++i) {
        const float* q = xyz + 3 * (size_t)(first + i);
        out[i] = make_float4(q[0], q[1], q[2], 0.0f);
    }
    return out;
}

int fl_grid_ok(const LusGrid& g) {
    return g.cell_size > 0.0f && g.gridX > 0 && g.gridY > 0 && g.gridZ > 0 &&
           (long long)g.gridX * g.gridY * g.gridZ < (1LL << 30);
}

}  // namespace

int lus_fluid_init(LusGpu* c, const LusFluidParams* p, int total, int sand_start, int sand_end, int solid_start,
                   int solid_end, const float* positions, const float* positions_star, const float* velocities,
                   int n_lambdas, const float* lambdas) {
    if (!c || !p || total <= 0 || !positions || !positions_star || !velocities || n_lambdas < 0 ||
        (n_lambdas && !lambdas))
        return lus::report_error(LUS_EINVAL, "lus_fluid_init: bad arguments");
    if (sand_start < 0 || sand_end < sand_start || sand_end > total || solid_start < 0 ||
        solid_end < solid_start || solid_end > total || (sand_end > sand_start && solid_end > solid_start &&
                                                         solid_start < sand_end && sand_start < solid_end))
        return lus::report_error(LUS_EINVAL, "lus_fluid_init: bad particle ranges");
    if (n_lambdas < sand_end)
        return lus::report_error(LUS_EINVAL,
                                 "lus_fluid_init: lambdas has %d entries, the moving slots reach %d (the "
                                 "reference writes lambdas[i] for every moving slot)",
                                 n_lambdas, sand_end);
    if (!fl_grid_ok(p->grid)) return lus::report_error(LUS_EINVAL, "lus_fluid_init: bad grid");
    if (!(p->kernelRadius > 0.0f) || !(p->rest_density != 0.0f) || !(p->mass > 0.0f))
        return lus::report_error(LUS_EINVAL, "lus_fluid_init: bad constants");
    const LusGrid& g = p->grid;
    const int C = g.gridX * g.gridY * g.gridZ;
    // solids never move: their cells are fixed, and must exist
    for (int i = solid_start; i < solid_end; ++i) {
        const float* q = positions_star + 3 * (size_t)i;
        const float px = q[0] / g.cell_size, py = q[1] / g.cell_size, pz = q[2] / g.cell_size;
        if (!(px >= 0.0f && py >= 0.0f && pz >= 0.0f) || (int)px >= g.gridX || (int)py >= g.gridY ||
            (int)pz >= g.gridZ)
            return lus::report_error(LUS_EINVAL, "lus_fluid_init: solid slot %d lies outside the grid", i);
    }
    const int dev = lus::ctx_device(c);
    hipStream_t s = lus::ctx_stream(c);
    FL_HIP(hipSetDevice(dev));
    FL_HIP(hipStreamSynchronize(s));
    lus::FluidState*& slot = lus::ctx_fluid(c);
    lus::fluid_free(slot);
    slot = nullptr;
    auto* F = new lus::FluidState();
    slot = F;
    F->p = *p;
    F->total = total;
    F->s0 = sand_start;
    F->s1 = sand_end;
    F->t0 = solid_start;
    F->t1 = solid_end;
    F->C = C;
    F->ns = solid_end - solid_start;
    const int n = sand_end - sand_start;
    lus::FluidK& k = F->k;
    k.r = p->particleRadius;
    k.h = p->kernelRadius;
    {
        volatile float hh = p->kernelRadius;
        k.h2 = hh * hh;
    }
    k.f = p->kernelFactor;
    k.ck = p->cubic_kernel_k;
    k.cl = p->cubic_kernel_l;
    k.mass = p->mass;
    k.rho0 = p->rest_density;
    k.eps = p->relaxation_epsilon;
    k.scorr_k = p->s_corr_k;
    k.scorr_n = p->s_corr_n;
    k.ipow = (p->s_corr_n >= 0.0f && p->s_corr_n <= 64.0f && p->s_corr_n == std::floor(p->s_corr_n))
                 ? (int)p->s_corr_n
                 : -1;
    for (int a = 0; a < 3; ++a) {
        volatile float gma = p->gravity[a] * p->mass;
        k.gm[a] = gma;
        volatile float hi = (float)(int)p->domain[a] - p->particleRadius;
        k.hi[a] = hi;
    }
    k.cs = g.cell_size;
    k.gX = g.gridX;
    k.gY = g.gridY;
    k.gZ = g.gridZ;
    k.gXZ = g.gridX * g.gridZ;
    k.s0 = sand_start;
    k.n = n;
    k.nl = n_lambdas;
    // W(0) and W(dq), the host evaluating the device's formula
    auto cubic = [&](float r) {
        volatile float q = (r * k.f) / k.h;
        if ((double)q > 1.0) return 0.0f;
        if ((double)q <= 0.5) {
            volatile float q2 = q * q;
            volatile float q3 = q2 * q;
            volatile float a = 6.0f * q3;
            volatile float b = 6.0f * q2;
            volatile float ab = a - b;
            volatile float ab1 = ab + 1.0f;
            return k.ck * ab1;
        }
        volatile float t = 1.0f - q;
        const double t3 = ((double)t * (double)t) * (double)t;
        volatile float two = 2.0f * (float)t3;
        return k.ck * two;
    };
    k.W0 = cubic(0.0f);
    k.Wdq = cubic(p->s_corr_dq);

    FL_HIP(fl_alloc(F, &F->x, total));
    FL_HIP(fl_alloc(F, &F->xs, total));
    FL_HIP(fl_alloc(F, &F->v, total));
    FL_HIP(fl_alloc(F, &F->lam, std::max(n_lambdas, 1)));
    FL_HIP(fl_alloc(F, &F->sorted, n));
    FL_HIP(fl_alloc(F, &F->key, n));
    FL_HIP(fl_alloc(F, &F->slot, n));
    FL_HIP(fl_alloc(F, &F->counts, C + 1));
    FL_HIP(fl_alloc(F, &F->cs, C + 1));
    FL_HIP(fl_alloc(F, &F->tmp, n));
    FL_HIP(fl_alloc(F, &F->perm, n));
    FL_HIP(fl_alloc(F, &F->ks, n));
    FL_HIP(fl_alloc(F, &F->tiles, lus::scan_tile_count(C + 1) + 1));
    FL_HIP(fl_alloc(F, &F->inv, n));
    FL_HIP(fl_alloc(F, &F->cnt, n + 1));
    FL_HIP(fl_alloc(F, &F->off, n + 1));
    FL_HIP(fl_alloc(F, &F->ltiles, lus::scan_tile_count(n + 1) + 1));
    FL_HIP(fl_alloc(F, &F->indeg, n));
    FL_HIP(fl_alloc(F, &F->q0, n));
    FL_HIP(fl_alloc(F, &F->q1, n));
    FL_HIP(fl_alloc(F, &F->ctl, lus::kCtlInts));
    FL_HIP(hipHostMalloc((void**)&F->host_ctl, sizeof(int) * lus::kCtlInts, hipHostMallocDefault));
    FL_HIP(hipMemsetAsync(F->counts, 0, sizeof(int) * (C + 1), s));

    std::vector<float4> h = aos_to_f4(positions, 0, total);
    FL_HIP(hipMemcpyAsync(F->x, h.data(), sizeof(float4) * total, hipMemcpyHostToDevice, s));
    FL_HIP(hipStreamSynchronize(s));
    h = aos_to_f4(positions_star, 0, total);
    FL_HIP(hipMemcpyAsync(F->xs, h.data(), sizeof(float4) * total, hipMemcpyHostToDevice, s));
    FL_HIP(hipStreamSynchronize(s));
    std::vector<float4> hv((size_t)total, make_float4(0.0f, 0.0f, 0.0f, 0.0f));
    for (int i = sand_start; i < sand_end; ++i) {
        const float* q = velocities + 3 * (size_t)i;
        hv[i] = make_float4(q[0], q[1], q[2], 0.0f);
    }
    FL_HIP(hipMemcpyAsync(F->v, hv.data(), sizeof(float4) * total, hipMemcpyHostToDevice, s));
    if (n_lambdas) FL_HIP(hipMemcpyAsync(F->lam, lambdas, sizeof(float) * n_lambdas, hipMemcpyHostToDevice, s));

    // solids: sorted once by cell (ascending slot inside a cell, as they
    // enter the reference's cell vectors)
    const int ns = F->ns;
    FL_HIP(fl_alloc(F, &F->osorted, ns));
    FL_HIP(fl_alloc(F, &F->ocs, C + 1));
    FL_HIP(fl_alloc(F, &F->oref, ns));
    {
        std::vector<int> okey((size_t)ns);
        for (int q = 0; q < ns; ++q) {
            const float* pp = positions_star + 3 * (size_t)(solid_start + q);
            okey[q] = ((int)(pp[1] / g.cell_size)) * g.gridX * g.gridZ + ((int)(pp[0] / g.cell_size)) * g.gridZ +
                      (int)(pp[2] / g.cell_size);
        }
        std::vector<int> ocnt((size_t)C + 1, 0), order((size_t)ns);
        for (int q = 0; q < ns; ++q) ocnt[okey[q] + 1]++;
        for (int cc = 0; cc < C; ++cc) ocnt[cc + 1] += ocnt[cc];
        std::vector<int> fillp(ocnt.begin(), ocnt.end() - 1);
        for (int q = 0; q < ns; ++q) order[fillp[okey[q]]++] = q;  // stable: ascending slot
        std::vector<float4> os((size_t)ns);
        std::vector<int> orf((size_t)ns);
        for (int q = 0; q < ns; ++q) {
            const float* pp = positions_star + 3 * (size_t)(solid_start + order[q]);
            os[q] = make_float4(pp[0], pp[1], pp[2], 0.0f);
            orf[q] = solid_start + order[q];
        }
        if (ns) {
            FL_HIP(hipMemcpyAsync(F->osorted, os.data(), sizeof(float4) * ns, hipMemcpyHostToDevice, s));
            FL_HIP(hipMemcpyAsync(F->oref, orf.data(), sizeof(int) * ns, hipMemcpyHostToDevice, s));
        }
        FL_HIP(hipMemcpyAsync(F->ocs, ocnt.data(), sizeof(int) * (C + 1), hipMemcpyHostToDevice, s));
        FL_HIP(hipStreamSynchronize(s));
    }
    F->stepped = false;
    return LUS_OK;
}

int lus_fluid_step(LusGpu* c, float dt) {
    if (!c) return lus::report_error(LUS_EINVAL, "lus_fluid_step: bad arguments");
    lus::FluidState* F = lus::ctx_fluid(c);
    if (!F) return lus::report_error(LUS_ESTATE, "lus_fluid_step: lus_fluid_init first");
    const int dev = lus::ctx_device(c);
    hipStream_t s = lus::ctx_stream(c);
    FL_HIP(hipSetDevice(dev));
    // dt = glm::clamp(dt, 0.001f, 0.01f) (Simulate.cpp:31)
    float t = dt < 0.001f ? 0.001f : dt;
    dt = 0.01f < t ? 0.01f : t;
    const lus::FluidK& k = F->k;
    const int n = k.n;
    FL_HIP(hipMemsetAsync(F->ctl, 0, sizeof(int) * lus::kCtlInts, s));
    F->last_levels = 0;
    F->last_entries = 0;
    if (n == 0) {
        F->stepped = true;
        return LUS_OK;
    }
    const int gb = lus::grid_for(n);
    lus::launch_kernel(lus::k_fl_predict, dim3(gb), dim3(lus::kBlock), 0, s, F->x, F->v, F->xs, k, dt, F->key,
                       F->C, F->ctl);
    FL_HIP(hipGetLastError());
    // stable counting sort of the moving particles by cell
    FL_HIP(lus::launch_count_keys(F->key, n, F->counts, F->slot, s));
    FL_HIP(lus::launch_exclusive_scan(F->counts, F->cs, F->C + 1, F->tiles, 1, s));
    FL_HIP(lus::launch_scatter(F->key, F->slot, F->cs, n, F->tmp, s));
    FL_HIP(lus::launch_stable_rank(F->tmp, F->key, F->cs, n, F->perm, F->ks, s));
    lus::launch_kernel(lus::k_fl_gather, dim3(gb), dim3(lus::kBlock), 0, s, F->xs, F->perm, k.s0, n, F->sorted,
                       F->inv);
    FL_HIP(hipGetLastError());
    // lists: counts, offsets, entries
    const float4* os = F->ns ? F->osorted : nullptr;
    lus::launch_kernel(lus::k_fl_lists, dim3(gb), dim3(lus::kBlock), 0, s, F->sorted, F->ks, F->cs, F->perm, F->inv,
                       os, F->ocs, F->oref, k, (const int*)nullptr, F->cnt, (int*)nullptr, 0);
    FL_HIP(hipGetLastError());
    FL_HIP(hipMemsetAsync(F->cnt + n, 0, sizeof(int), s));
    FL_HIP(lus::launch_exclusive_scan(F->cnt, F->off, n + 1, F->ltiles, 0, s));
    FL_HIP(hipMemcpyAsync(F->host_ctl, F->off + n, sizeof(int), hipMemcpyDeviceToHost, s));
    FL_HIP(hipMemcpyAsync(F->host_ctl + 1, F->ctl + lus::kCtlErr, sizeof(int), hipMemcpyDeviceToHost, s));
    FL_HIP(hipStreamSynchronize(s));
    if (F->host_ctl[1] & lus::kErrKey)
        return lus::report_error(LUS_EINVAL,
                                 "lus_fluid_step: a predicted position lies outside the grid (the reference "
                                 "indexes outside its cell vector there)");
    const long total = F->host_ctl[0];
    if (total < 0) return lus::report_error(LUS_EINVAL, "lus_fluid_step: more than 2^31 list entries");
    if (total > F->cap) {
        if (F->idx_mem) (void)hipFree(F->idx_mem);
        F->idx_mem = nullptr;
        F->cap = 0;
        const long want = total + total / 4 + 1024;
        FL_HIP(hipMalloc(&F->idx_mem, sizeof(int) * (size_t)want));
        F->cap = want;
    }
    F->idx = static_cast<int*>(F->idx_mem);
    F->last_entries = total;
    lus::launch_kernel(lus::k_fl_lists, dim3(gb), dim3(lus::kBlock), 0, s, F->sorted, F->ks, F->cs, F->perm, F->inv,
                       os, F->ocs, F->oref, k, (const int*)F->off, F->cnt, F->idx, 1);
    FL_HIP(hipGetLastError());
    lus::launch_kernel(lus::k_fl_lambda, dim3(gb), dim3(lus::kBlock), 0, s, (const float4*)F->xs,
                       (const int*)F->off, (const int*)F->idx, k, F->lam, F->indeg, F->q0, F->ctl);
    FL_HIP(hipGetLastError());
    // the sweep: one launch per dependency level, checked every kBatch levels
    constexpr int kBatch = 32;
    const int sweep_grid = std::min(lus::grid_for(n), 64);
    int level = 0, done_prev = -1;
    for (;;) {
        for (int b = 0; b < kBatch; ++b, ++level) {
            int* in = (level % 2 == 0) ? F->q0 : F->q1;
            int* out = (level % 2 == 0) ? F->q1 : F->q0;
            lus::launch_kernel(lus::k_fl_sweep, dim3(sweep_grid), dim3(lus::kBlock), 0, s, F->x, F->xs, F->v,
                               (const float*)F->lam, (const int*)F->off, (const int*)F->idx, k, dt, level,
                               (const int*)in, out, F->indeg, F->ctl);
        }
        FL_HIP(hipGetLastError());
        FL_HIP(hipMemcpyAsync(F->host_ctl, F->ctl, sizeof(int) * lus::kCtlInts, hipMemcpyDeviceToHost, s));
        FL_HIP(hipStreamSynchronize(s));
        const int done = F->host_ctl[lus::kCtlDone];
        if (done >= n) break;
        if (done == done_prev)
            return lus::report_error(LUS_EHIP, "lus_fluid_step: the sweep stopped after %d of %d particles", done,
                                     n);
        done_prev = done;
    }
    F->last_levels = F->host_ctl[lus::kCtlLevels];
    if (F->host_ctl[lus::kCtlErr] & lus::kErrLambda)
        return lus::report_error(LUS_EINVAL,
                                 "lus_fluid_step: a neighbour list is longer than the lambdas array (the "
                                 "reference reads past its vector there)");
    F->stepped = true;
    return LUS_OK;
}

int lus_fluid_download(LusGpu* c, float* positions, float* positions_star, float* velocities, float* lambdas) {
    if (!c) return lus::report_error(LUS_EINVAL, "lus_fluid_download: bad arguments");
    lus::FluidState* F = lus::ctx_fluid(c);
    if (!F) return lus::report_error(LUS_ESTATE, "lus_fluid_download: lus_fluid_init first");
    hipStream_t s = lus::ctx_stream(c);
    FL_HIP(hipSetDevice(lus::ctx_device(c)));
    std::vector<float4> h((size_t)F->total);
    auto pull = [&](const float4* src, float* dst, int a, int b) -> hipError_t {
        if (!dst) return hipSuccess;
        hipError_t e = hipMemcpyAsync(h.data(), src, sizeof(float4) * F->total, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) return e;
        for (int i = a; i < b; ++i) {
            dst[3 * (size_t)i] = h[i].x;
            dst[3 * (size_t)i + 1] = h[i].y;
            dst[3 * (size_t)i + 2] = h[i].z;
        }
        return hipSuccess;
    };
    FL_HIP(pull(F->x, positions, 0, F->total));
    FL_HIP(pull(F->xs, positions_star, 0, F->total));
    FL_HIP(pull(F->v, velocities, F->s0, F->s1));
    if (lambdas && F->k.nl) {
        FL_HIP(hipMemcpyAsync(lambdas, F->lam, sizeof(float) * F->k.nl, hipMemcpyDeviceToHost, s));
        FL_HIP(hipStreamSynchronize(s));
    }
    return LUS_OK;
}

int lus_fluid_stats(LusGpu* c, int* levels, long* list_entries) {
    if (!c) return lus::report_error(LUS_EINVAL, "lus_fluid_stats: bad arguments");
    lus::FluidState* F = lus::ctx_fluid(c);
    if (!F) return lus::report_error(LUS_ESTATE, "lus_fluid_stats: lus_fluid_init first");
    if (levels) *levels = F->last_levels;
    if (list_entries) *list_entries = F->last_entries;
    return LUS_OK;
}
