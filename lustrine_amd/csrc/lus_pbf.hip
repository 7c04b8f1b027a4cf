// lus_pbf.hip -- gfx950 kernels of the position-based-fluids substep.
//
// Per step (arrays in cell order after the counting sort, one float4 per
// particle):
//   k_pbf_key      predict x* = clamp(x + (v + g dt) dt), cell key, histogram
//   (scan, scatter: lus_kernels.hip)
//   k_pbf_permute  stable cell order; writes x (old), x* (predicted), id, key
//   k_pbf_lambda   first call of the step also builds the neighbour sets
//   k_pbf_deltap   Jacobi position correction with s_corr, box clamp
//   k_pbf_xsph     XSPH viscosity, commit
//
// Neighbour sets.  N(i) = { j : |x*_i - x*_j|^2 <= h^2 } from the predicted
// positions, once per step.  The 27-cell stencil of a particle is 9 rows of
// contiguous slots (y-major cells, z contiguous).  The first lambda pass tests
// every candidate once, row by row through a 64-bit hit mask per row, visits
// the hits and appends them to i's neighbour list, in ascending slot order
// (the order of the CPU oracle's lists).  The other passes of the step walk
// the lists, one entry per iteration in every lane, so a wave runs
// max_i |N(i)| iterations instead of the sum over the 9 rows of the longest
// row (about half as many on a dam break).  A particle with more than
// kNbrCap neighbours (or with an entry beyond 16 bits) re-tests its candidates against the stored predicted
// positions in every pass, which gives the same set in the same order.  The
// tiled passes keep the row masks instead (masks[r * n + i], bit t = slot
// j0_r + t) and flag a particle whose stencil has a row longer than 64 slots
// in its key (kOvf).
//
// Arithmetic modes (template kFast, chosen by LusPbfParams::exact):
//   exact  every expression in the oracle's order, IEEE '/' and sqrtf, no
//          FMA: bit-identical to the CPU oracle;
//   fast   fused multiply-adds and the hardware reciprocal square root
//          (v_rsq_f32) replace sqrt + divide in the neighbour loops.
// The neighbour test follows the mode: exact, r2 = dx dx + dy dy + dz dz rounded
// at every step and r2 <= h2; fast, h2 - r2 by three fused multiply-adds and a
// hit where it is >= 0 (the two differ only for pairs within rounding of r = h,
// where every kernel weight and gradient vanishes).
#include <hip/hip_runtime.h>

#include <climits>

#include "lus_device.hpp"
#include "lus_internal.hpp"
#include "lus_launch.hpp"

namespace lus {

constexpr int kOvf = 1 << 30;  // keys are < 2^30 (checked at init)

// the small-scene list build and walks (kSubLanes lanes per particle), defined below
hipError_t launch_lambda_build_sub(const float4* xp, const int* key2, const int* cs, int n, PbfK k, float4* xl,
                                   int* maxc, unsigned long long* masks, const int* gate, hipStream_t s);
hipError_t launch_lambda_sub(const float4* xa, int* key2, const int* cs, int n, PbfK k, float4* xl, int* maxc,
                             unsigned long long* masks, const float4* xp, const int* gate, hipStream_t s);
hipError_t launch_deltap_sub(const float4* xl, int* key2, const int* cs, int n, PbfK k, float4* xa,
                             const float4* xold, float4* vout, float dt, unsigned long long* masks,
                             const float4* xp, const int* gate, hipStream_t s);
hipError_t launch_xsph_sub(const float4* xa, const float4* vt, int* key2, const int* cs, int n, PbfK k,
                           float4* vout, float4* xout, unsigned long long* masks, const float4* xp,
                           const int* gate, hipStream_t s,
                           const NextKeys& nk);

// clamp box and cell y offset of a particle: the scene table entry of a batch,
// or the single scene's bounds
struct Bounds {
    float lo[3], hi[3];
    int yoff;
};

__device__ __forceinline__ Bounds bounds_of(const PbfK& k, int s) {
    Bounds b;
    if (k.scenes) {
        const SceneBox sb = k.scenes[s];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            b.lo[a] = sb.lo[a];
            b.hi[a] = sb.hi[a];
        }
        b.yoff = sb.yoff;
    } else {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            b.lo[a] = k.lo[a];
            b.hi[a] = k.hi[a];
        }
        b.yoff = 0;
    }
    return b;
}

__device__ __forceinline__ float4 pbf_predict(float4 x, float4 v, const PbfK& k, const Bounds& b,
                                              float dt) {
    float vx = v.x + k.g[0] * dt;
    float vy = v.y + k.g[1] * dt;
    float vz = v.z + k.g[2] * dt;
    float4 r;
    r.x = rclamp(x.x + vx * dt, b.lo[0], b.hi[0]);
    r.y = rclamp(x.y + vy * dt, b.lo[1], b.hi[1]);
    r.z = rclamp(x.z + vz * dt, b.lo[2], b.hi[2]);
    r.w = 0.0f;
    return r;
}

// get_cell_id (src/neighbors/Utils.hpp:24-33) in the (batch) grid: the y cell
// index is shifted by the scene's offset; c: the phase run of a mixed scene
__device__ __forceinline__ int pbf_cell(float4 p, const PbfK& k, int yoff, int c) {
    const float px = p.x / k.cs;
    const float py = p.y / k.cs;
    const float pz = p.z / k.cs;
    return ((((int)py + yoff) * k.gX + (int)px) * k.nph + c) * k.gZ + (int)pz;
}

// cell column (cy, cx) and cz of a key ((cy gX + cx) nph + c) gZ + cz
__device__ __forceinline__ void key_cell(int key, const PbfK& k, int& cy, int& cx, int& cz) {
    const int q = key / k.gZ;
    cz = key - q * k.gZ;
    const int col = k.nph == 1 ? q : q / k.nph;
    cy = col / k.gX;
    cx = col - cy * k.gX;
}

// first key of phase run c of cell column (y, x) (z = 0)
__device__ __forceinline__ int row_key(int y, int x, int c, const PbfK& k) {
    return ((y * k.gX + x) * k.nph + c) * k.gZ;
}

// h2 - r2 of the fast mode's neighbour test: three fused multiply-adds
__device__ __forceinline__ float h2_less_r2_fast(float dx, float dy, float dz, const PbfK& k) {
    return __builtin_fmaf(-dz, dz, __builtin_fmaf(-dy, dy, __builtin_fmaf(-dx, dx, k.h2)));
}

// the neighbour test of the mode kFast: exact (no FMA, oracle order: r2 <=
// h2) or fast (h2 - r2 >= 0 by fused multiply-adds)
template <bool kFast = false>
__device__ __forceinline__ bool within_h(const float4 a, const float4 b, const PbfK& k) {
    const float dx = a.x - b.x;
    const float dy = a.y - b.y;
    const float dz = a.z - b.z;
    if constexpr (kFast) return h2_less_r2_fast(dx, dy, dz, k) >= 0.0f;
    const float r2 = dx * dx + dy * dy + dz * dz;
    return r2 <= k.h2;
}
template <bool kFast = false>
__device__ __forceinline__ bool within_h(const float4 a, const float3 b, const PbfK& k) {
    return within_h<kFast>(a, make_float4(b.x, b.y, b.z, 0.0f), k);
}
// the neighbour test of the scene's mode (k.exact), for loops that are not
// specialised on it
__device__ __forceinline__ bool within_h_k(const float4 a, const float4 b, const PbfK& k) {
    return k.exact ? within_h<false>(a, b, k) : within_h<true>(a, b, k);
}
// x, y, z of slot t of a particle array (one 12-byte load: a candidate test
// reads no more)
__device__ __forceinline__ float3 xyz_at(const float4* __restrict__ a, int t) {
    return *reinterpret_cast<const float3*>(a + t);
}

__device__ __forceinline__ int low_bit(unsigned int m) { return __builtin_ctz(m); }
__device__ __forceinline__ int low_bit(unsigned long long m) { return __builtin_ctzll(m); }

// aux[j], or zeros without aux.  Both arms are values: a conditional between
// aux[j] and a named zero vector would select between two addresses and put
// the zero vector in scratch memory.
__device__ __forceinline__ float4 aux_at(const float4* __restrict__ aux, int j) {
    return aux ? aux[j] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// Visit the set bits of one row mask in ascending order, calling
// f(j, cur[j + delta], aux[j + delta]) for slot j = base + bit; the next
// neighbour's data is loaded before f runs on the current one.  delta maps a
// slot to its copy in a staged (LDS) array, 0 for the global arrays.
template <typename M, class F>
__device__ __forceinline__ void visit_row(M mm, int base, int delta, const float4* __restrict__ cur,
                                          const float4* __restrict__ aux, F& f) {
    if (!mm) return;
    int j = base + low_bit(mm);
    mm &= mm - (M)1;
    float4 pa = cur[j + delta];
    float4 aa = aux_at(aux, j + delta);
    while (mm) {
        const int jn = base + low_bit(mm);
        mm &= mm - (M)1;
        const float4 pb = cur[jn + delta];
        const float4 ab = aux_at(aux, jn + delta);
        f(j, pa, aa);
        j = jn;
        pa = pb;
        aa = ab;
    }
    f(j, pa, aa);
}

// float4 of slot j of a particle array, addressed with a 32-bit byte offset
// from the array base (a context holds at most kMaxParticles < 2^28 slots)
__device__ __forceinline__ float4 at4(const float4* __restrict__ a, int j) {
    return *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(a) + (unsigned)j * 16u);
}

// Neighbour-list buffer (layout in lus_internal.hpp): headers, then int4
// groups of eight 16-bit entries.
__device__ __forceinline__ int4* nbr_hdr(unsigned long long* nb) { return reinterpret_cast<int4*>(nb); }
// group 0 of slot i's list; group g is 64 int4 further
__device__ __forceinline__ int4* nbr_list(unsigned long long* nb, int n, int i) {
    int4* lists = reinterpret_cast<int4*>(nb) + (size_t)n;
    return lists + (size_t)(i >> 6) * (kNbrGroups * 64) + (i & 63);
}

// Appends the neighbours of slot i to its list while a build visits them in
// ascending slot order: entries relative to the first neighbour of their
// layer (stencil rows 0-2, 3-5, 6-8), eight 16-bit entries per int4 group,
// one 16-byte store per completed group (lst: group 0 of slot i's list, group
// g 64 int4 further); entries past kNbrCap are counted, not stored.
struct ListWriter {
    int c = 0;
    int lb = -1;  // first neighbour of the current layer (-1: none yet)
    int jl = -1;  // last neighbour announced by row_hits in the current layer
    int b0, b1;
    // min(n0, 255) | min(n01, 255) << 8, bit 16: an entry does not fit 16 bits
    int meta = 0;
    unsigned int q0 = 0u, q1 = 0u, q2 = 0u, q3 = 0u;  // the group being filled, entry 0 lowest

    __device__ __forceinline__ explicit ListWriter(int i) : b0(i), b1(i) {}
    __device__ __forceinline__ void push(unsigned int e) {
        q0 = (q0 >> 16) | (q1 << 16);
        q1 = (q1 >> 16) | (q2 << 16);
        q2 = (q2 >> 16) | (q3 << 16);
        q3 = (q3 >> 16) | (e << 16);
    }
    // at the top of stencil row r: close layer 0 (r = 3) or layer 1 (r = 6)
    // (selects, not branches: branches on r let the compiler keep b0 and b1
    // in a stack slot indexed by the layer)
    __device__ __forceinline__ void row(int r, int i) {
        const int first = lb < 0 ? i : lb;
        const int cc = c < 255 ? c : 255;
        const bool close = r == 3 || r == 6;
        meta |= r == 3 ? cc : (r == 6 ? cc << 8 : 0);
        meta |= (close && jl - first > 0xFFFF) ? 0x10000 : 0;
        b0 = r == 3 ? first : b0;
        b1 = r == 6 ? first : b1;
        lb = close ? -1 : lb;
        jl = close ? -1 : jl;
    }
    // A row whose hits run from jfirst to jlast, appended next with
    // append_row: the layer's first neighbour is set here, and its entries
    // (ascending within the layer) are checked against 16 bits once, through
    // the layer's last one, when the layer closes.
    __device__ __forceinline__ void row_hits(int jfirst, int jlast) {
        lb = lb < 0 ? jfirst : lb;
        jl = jlast;
    }
    __device__ __forceinline__ bool append_row(int j) {
        push((unsigned int)(j - lb) & 0xFFFFu);
        ++c;
        return (c & 7) == 0 && c <= kNbrCap;
    }
    // append j; true when that completes a group to store (group(), number
    // (c >> 3) - 1)
    __device__ __forceinline__ bool append(int j) {
        lb = lb < 0 ? j : lb;
        const unsigned int e = (unsigned int)(j - lb);
        meta |= e > 0xFFFFu ? 0x10000 : 0;
        push(e & 0xFFFFu);
        ++c;
        return (c & 7) == 0 && c <= kNbrCap;
    }
    __device__ __forceinline__ int4 group() const { return make_int4((int)q0, (int)q1, (int)q2, (int)q3); }
    __device__ __forceinline__ void emit(int j, int4* __restrict__ lst) {
        if (append(j)) lst[(size_t)((c >> 3) - 1) * 64] = group();
    }
    // the last, partial group (zero entries in the unused places: they decode
    // to b2, a particle's slot, loaded by a walk but not visited) and the
    // header hdr[i]
    __device__ __forceinline__ void finish(int i, int4* __restrict__ hdr, int4* __restrict__ lst) {
        const int b2 = lb < 0 ? i : lb;
        meta |= jl - b2 > 0xFFFF ? 0x10000 : 0;
        const int rem = c & 7;
        if (rem != 0 && c < kNbrCap) {
            for (int u = rem; u < 8; ++u) push(0u);
            lst[(size_t)(c >> 3) * 64] = make_int4((int)q0, (int)q1, (int)q2, (int)q3);
        }
        // c <= kNbrCap: n0 and n01 are below 255
        const int w = (c > kNbrCap || (meta & 0x10000)) ? kNbrRetest : (c | (meta & 0xFFFF) << 16);
        hdr[i] = make_int4(b0, b1, b2, w);
    }
};

// Test the 27-cell candidates of slot i (positions xp) row by row, in
// ascending slot order -- a row of at most 64 slots through a 64-bit hit mask
// (32-bit mask arithmetic when no lane of the wave has a hit above bit 31), a
// longer row slot by slot -- and call f(j, cur[j], aux[j]) on every hit while
// appending it to i's list (ListWriter; the lanes of a wave write whole
// segments of a group).  Writes the header hdr[i].
template <bool kFast, class F>
__device__ __forceinline__ void build_list_t(int i, const int* __restrict__ key2, const int* __restrict__ cs,
                                             const float4* __restrict__ xp, const PbfK& k, int4* __restrict__ hdr,
                                             int4* __restrict__ lst, const float4* __restrict__ cur,
                                             const float4* __restrict__ aux, F& f) {
    int cy, cx, cz;
    key_cell(key2[i] & ~kOvf, k, cy, cx, cz);
    const int z0 = cz > 0 ? cz - 1 : 0;
    const int z1 = cz + 1 < k.gZ ? cz + 1 : k.gZ - 1;
    const float4 pi = xp[i];
    ListWriter lw(i);
    auto emit = [&](int j, const float4 pj, const float4 aj) {
        lw.emit(j, lst);
        f(j, pj, aj);
    };
#pragma unroll 1
    for (int r = 0; r < 9; ++r) {
        lw.row(r, i);
        const int y = cy + r / 3 - 1;
        const int x = cx + r % 3 - 1;
        if (y < 0 || y >= k.gY || x < 0 || x >= k.gX) continue;
#pragma unroll 1
        for (int c = 0; c < k.nph; ++c) {
            const int row = row_key(y, x, c, k);
            const int j0 = cs[row + z0];
            const int j1 = cs[row + z1 + 1];
            const int L = j1 - j0;
            if (L > 64) {
                for (int j = j0; j < j1; ++j)
                    if (within_h<kFast>(pi, xp[j], k)) emit(j, at4(cur, j), aux_at(aux, j));
                continue;
            }
            unsigned long long mm = 0ull;
            const float4* __restrict__ base = xp + j0;
            int t = 0;
            for (; t + 4 <= L; t += 4) {
                const float3 c0 = xyz_at(base, t), c1 = xyz_at(base, t + 1), c2 = xyz_at(base, t + 2),
                             c3 = xyz_at(base, t + 3);
                mm |= within_h<kFast>(pi, c0, k) ? (1ull << t) : 0ull;
                mm |= within_h<kFast>(pi, c1, k) ? (2ull << t) : 0ull;
                mm |= within_h<kFast>(pi, c2, k) ? (4ull << t) : 0ull;
                mm |= within_h<kFast>(pi, c3, k) ? (8ull << t) : 0ull;
            }
            for (; t < L; ++t) mm |= within_h<kFast>(pi, xyz_at(base, t), k) ? (1ull << t) : 0ull;
            if (__any((int)((mm >> 32) != 0ull)))
                visit_row<unsigned long long>(mm, j0, 0, cur, aux, emit);
            else
                visit_row<unsigned int>((unsigned int)mm, j0, 0, cur, aux, emit);
        }
    }
    lw.finish(i, hdr, lst);
}

template <class F>
__device__ __forceinline__ void build_list(int i, const int* __restrict__ key2, const int* __restrict__ cs,
                                           const float4* __restrict__ xp, const PbfK& k, int4* __restrict__ hdr,
                                           int4* __restrict__ lst, const float4* __restrict__ cur,
                                           const float4* __restrict__ aux, F& f) {
    if (k.exact)
        build_list_t<false>(i, key2, cs, xp, k, hdr, lst, cur, aux, f);
    else
        build_list_t<true>(i, key2, cs, xp, k, hdr, lst, cur, aux, f);
}

// Walk forms of visit_list: kWalkWide, one group of eight entries per step
// with all eight gathers in flight (the plain lambda and delta-p passes: one
// gather per neighbour, within 64 VGPRs); kWalkPairs, one group per step in
// two halves of four (the XSPH passes of both scenes: two gathers per
// neighbour); kWalkHalf, one half per loop step (the mixed scene's lambda and
// correction passes, whose per-neighbour bodies are long: four copies of the
// body instead of eight).
constexpr int kWalkHalf = 0, kWalkPairs = 1, kWalkWide = 2;

// Walk the list of header hd (|N(i)| <= kNbrCap), calling f(j, cur[j],
// aux[j]) in list order.  A step's neighbour data are requested, then the
// next group's entries (so waiting for the data never waits for the list),
// then f runs.  Every lane of a wave is in the same group, so a group load is
// one whole 1 KB segment.  Places of the last group past |N(i)| decode to
// slot b2 (in-bounds data, not visited); the next group is used only when the
// walk goes on.
template <int kWalk, class F>
__device__ __forceinline__ void visit_list(const int4 hd, const int4* __restrict__ lst,
                                           const float4* __restrict__ cur, const float4* __restrict__ aux,
                                           F& f) {
    const int c = hd.w & 0xFFFF;
    if (c <= 0) return;
    const int n0 = (hd.w >> 16) & 0xFF, n01 = (int)((unsigned int)hd.w >> 24);
    // slot of entry t holding e
    auto slot = [&](int t, unsigned int e) { return (t < n0 ? hd.x : (t < n01 ? hd.y : hd.z)) + (int)e; };
    const float4 z = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    // four entries from the two words wa (t, t + 1) and wb (t + 2, t + 3)
    auto visit4 = [&](int t, unsigned int wa, unsigned int wb, auto&& between) {
        const int j0 = slot(t, wa & 0xFFFFu), j1 = slot(t + 1, wa >> 16);
        const int j2 = slot(t + 2, wb & 0xFFFFu), j3 = slot(t + 3, wb >> 16);
        const float4 p0 = at4(cur, j0), p1 = at4(cur, j1), p2 = at4(cur, j2), p3 = at4(cur, j3);
        float4 a0 = z, a1 = z, a2 = z, a3 = z;
        if (aux) {
            a0 = at4(aux, j0);
            a1 = at4(aux, j1);
            a2 = at4(aux, j2);
            a3 = at4(aux, j3);
        }
        between();
        // keep every component of the first neighbour in the step's loads
        // (a component used only under f's branch would be loaded there,
        // after the next group's entries, and wait for them)
        asm volatile("" ::"v"(p0.x), "v"(p0.y), "v"(p0.z), "v"(p0.w));
        f(j0, p0, a0);
        if (t + 1 < c) f(j1, p1, a1);
        if (t + 2 < c) f(j2, p2, a2);
        if (t + 3 < c) f(j3, p3, a3);
    };
    int4 g = lst[0];
    if constexpr (kWalk == kWalkHalf) {
        int4 gn = g;
        for (int t = 0; t < c; t += 4) {
            const bool lo = (t & 4) == 0;  // uniform
            const int gn_i = (t >> 3) + 1 < kNbrGroups ? (t >> 3) + 1 : kNbrGroups - 1;
            visit4(t, (unsigned int)(lo ? g.x : g.z), (unsigned int)(lo ? g.y : g.w), [&] {
                if (lo) gn = lst[(size_t)gn_i * 64];
            });
            if (!lo) g = gn;
        }
        return;
    }
    for (int t = 0; t < c; t += 8) {
        const int gn_i = (t >> 3) + 1 < kNbrGroups ? (t >> 3) + 1 : kNbrGroups - 1;
        const unsigned int w0 = (unsigned int)g.x, w1 = (unsigned int)g.y;
        const unsigned int w2 = (unsigned int)g.z, w3 = (unsigned int)g.w;
        int4 gn;
        if constexpr (kWalk == kWalkWide) {
            const int j0 = slot(t, w0 & 0xFFFFu), j1 = slot(t + 1, w0 >> 16);
            const int j2 = slot(t + 2, w1 & 0xFFFFu), j3 = slot(t + 3, w1 >> 16);
            const int j4 = slot(t + 4, w2 & 0xFFFFu), j5 = slot(t + 5, w2 >> 16);
            const int j6 = slot(t + 6, w3 & 0xFFFFu), j7 = slot(t + 7, w3 >> 16);
            const float4 p0 = at4(cur, j0), p1 = at4(cur, j1), p2 = at4(cur, j2), p3 = at4(cur, j3);
            const float4 p4 = at4(cur, j4), p5 = at4(cur, j5), p6 = at4(cur, j6), p7 = at4(cur, j7);
            gn = lst[(size_t)gn_i * 64];
            asm volatile("" ::"v"(p0.x), "v"(p0.y), "v"(p0.z), "v"(p0.w));
            asm volatile("" ::"v"(p4.x), "v"(p4.y), "v"(p4.z), "v"(p4.w));
            f(j0, p0, z);
            if (t + 1 < c) f(j1, p1, z);
            if (t + 2 < c) f(j2, p2, z);
            if (t + 3 < c) f(j3, p3, z);
            if (t + 4 < c) f(j4, p4, z);
            if (t + 5 < c) f(j5, p5, z);
            if (t + 6 < c) f(j6, p6, z);
            if (t + 7 < c) f(j7, p7, z);
        } else {
            visit4(t, w0, w1, [&] { gn = lst[(size_t)gn_i * 64]; });
            if (t + 4 < c) visit4(t + 4, w2, w3, [] {});
        }
        g = gn;
    }
}

// Visit N(i) in ascending slot order, calling f(j, cur[j], aux[j]).  build:
// test the candidates, visiting and listing the hits (build_list); otherwise
// walk the stored list.  A particle whose list overflowed re-tests its candidates
// against the stored predicted positions xp (the same set in the same order).
// nb: the neighbour-list buffer.
template <bool kBuild, int kWalk = kWalkHalf, class F>
__device__ __forceinline__ void for_each_neighbor(int i, int n, int* __restrict__ key2,
                                                  const int* __restrict__ cs,
                                                  unsigned long long* __restrict__ nb,
                                                  const float4* __restrict__ xp,
                                                  const float4* __restrict__ cur,
                                                  const float4* __restrict__ aux, const PbfK& k,
                                                  F&& f) {
    int4* __restrict__ lst = nbr_list(nb, n, i);
    if (kBuild) {
        build_list(i, key2, cs, xp, k, nbr_hdr(nb), lst, cur, aux, f);
        return;
    }
    const int4 hd = nbr_hdr(nb)[i];
    if ((hd.w & 0xFFFF) <= kNbrCap) {
        visit_list<kWalk>(hd, lst, cur, aux, f);
        return;
    }
    int cy, cx, cz;
    key_cell(key2[i] & ~kOvf, k, cy, cx, cz);
    const int z0 = cz > 0 ? cz - 1 : 0;
    const int z1 = cz + 1 < k.gZ ? cz + 1 : k.gZ - 1;
    const float4 pi = xp[i];
#pragma unroll 1
    for (int r = 0; r < 9; ++r) {
        const int y = cy + r / 3 - 1;
        const int x = cx + r % 3 - 1;
        if (y < 0 || y >= k.gY || x < 0 || x >= k.gX) continue;
        for (int c = 0; c < k.nph; ++c) {
            const int row = row_key(y, x, c, k);
            const int j0 = cs[row + z0];
            const int j1 = cs[row + z1 + 1];
            for (int j = j0; j < j1; ++j)
                if (within_h_k(pi, xp[j], k)) f(j, cur[j], aux_at(aux, j));
        }
    }
}

// ---------------------------------------------------------------------------
// Tiled neighbour passes.  A tile is at most kTileT consecutive slots of one
// (y, x) cell column, cells z in [za, zb].  Every stencil row of its particles
// lies in the 9 rows (y +- 1, x +- 1) over cells [za - 1, zb + 1]: the block
// stages those 9 slot ranges of the pass's input (and their cell starts) in
// LDS once, and the particles read their neighbours from there.  A tile whose
// staged ranges exceed the LDS capacity reads from global memory instead.
// The visiting order and the arithmetic are those of for_each_neighbor, so
// both paths give the same bits.
// ---------------------------------------------------------------------------
struct TileShared {
    int cs[9][kTileZMax + 3];  // cell starts of row r, cells zlo .. zhi + 1
    int rb[9];                 // first staged slot of row r
    int lo[10];                // LDS offset of row r's copy; lo[9] = staged count
    int y, x, zlo, zhi, i0, i1, staged;
};

// XCD-aware first tile of this block; tiles advance by gridDim.x (a multiple of 8)
__device__ __forceinline__ int tile_first() {
    const int G = gridDim.x, b = blockIdx.x;
    return (b % kXcd) * (G / kXcd) + b / kXcd;
}

// Stage tile `tl` (all threads of the block call it; ends with a barrier).
__device__ __forceinline__ void tile_stage(TileShared& S, int4 tl, const int* __restrict__ cs,
                                           const PbfK& k, const float4* __restrict__ cur,
                                           const float4* __restrict__ aux, float4* __restrict__ lds,
                                           float4* __restrict__ lds_aux, int lcap) {
    const int tid = threadIdx.x;
    if (tid == 0) {
        const int y = tl.z / k.gXZ;
        const int x = (tl.z - y * k.gXZ) / k.gZ;
        const int za = tl.w & 0xFFFF, zb = tl.w >> 16;
        S.i0 = tl.x;
        S.i1 = tl.y;
        S.y = y;
        S.x = x;
        S.zlo = za > 0 ? za - 1 : 0;
        S.zhi = zb + 1 < k.gZ ? zb + 1 : k.gZ - 1;
    }
    __syncthreads();
    const int w = S.zhi - S.zlo + 2;
    for (int t = tid; t < 9 * w; t += blockDim.x) {
        const int r = t / w, q = t - r * w;
        const int yy = S.y + r / 3 - 1, xx = S.x + r % 3 - 1;
        const bool valid = yy >= 0 && yy < k.gY && xx >= 0 && xx < k.gX;
        S.cs[r][q] = valid ? cs[yy * k.gXZ + xx * k.gZ + S.zlo + q] : 0;
    }
    __syncthreads();
    if (tid == 0) {
        int acc = 0;
        for (int r = 0; r < 9; ++r) {
            S.rb[r] = S.cs[r][0];
            S.lo[r] = acc;
            acc += S.cs[r][w - 1] - S.cs[r][0];
        }
        S.lo[9] = acc;
        S.staged = acc <= lcap ? 1 : 0;
    }
    __syncthreads();
    if (S.staged) {
        const int tot = S.lo[9];
        for (int t = tid; t < tot; t += blockDim.x) {
            int r = 0;
            while (r < 8 && t >= S.lo[r + 1]) ++r;
            const int j = S.rb[r] + (t - S.lo[r]);
            lds[t] = cur[j];
            if (aux) lds_aux[t] = aux[j];
        }
    }
    __syncthreads();
}

// for_each_neighbor for particle i of the staged tile S.  kStaged: read the
// neighbours from the LDS copies (src = lds, delta per row), otherwise from
// the global arrays (src = cur, delta 0); separate instantiations keep the
// LDS reads in the LDS address space.
template <bool kBuild, bool kStaged, class F>
__device__ __forceinline__ void for_each_neighbor_tiled_impl(const TileShared& S, int i, int n,
                                                             int* __restrict__ key2,
                                                             unsigned long long* __restrict__ masks,
                                                             const float4* __restrict__ xp,
                                                             const float4* __restrict__ src,
                                                             const float4* __restrict__ src_aux, const PbfK& k,
                                                             F& f) {
    const int kw = key2[i];
    const int key = kw & ~kOvf;
    const int cz = key - (key / k.gZ) * k.gZ;
    const int q0 = (cz > 0 ? cz - 1 : 0) - S.zlo;
    const int q1 = (cz + 1 < k.gZ ? cz + 1 : k.gZ - 1) + 1 - S.zlo;
    bool ovf;
    if (kBuild) {
        int maxlen = 0;
#pragma unroll
        for (int r = 0; r < 9; ++r) {
            const int l = S.cs[r][q1] - S.cs[r][q0];  // 0 for rows outside the grid
            maxlen = l > maxlen ? l : maxlen;
        }
        ovf = maxlen > 64;
        if (ovf) key2[i] = kw | kOvf;
    } else {
        ovf = (kw & kOvf) != 0;
    }
    const float4 pi = (kBuild || ovf) ? xp[i] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll 1
    for (int r = 0; r < 9; ++r) {
        const int y = S.y + r / 3 - 1;
        const int x = S.x + r % 3 - 1;
        if (y < 0 || y >= k.gY || x < 0 || x >= k.gX) continue;
        const int j0 = S.cs[r][q0];
        const int j1 = S.cs[r][q1];
        const int delta = kStaged ? S.lo[r] - S.rb[r] : 0;
        if (ovf) {
            for (int j = j0; j < j1; ++j)
                if (within_h_k(pi, xp[j], k)) f(j, src[j + delta], aux_at(src_aux, j + delta));
            continue;
        }
        unsigned long long mm;
        if (kBuild) {
            // candidates from the staged copy of x* (build passes stage x*)
            mm = 0ull;
            const float4* __restrict__ base = src + j0 + delta;
            const int L = j1 - j0;
            int t = 0;
            for (; t + 4 <= L; t += 4) {
                const float4 c0 = base[t], c1 = base[t + 1], c2 = base[t + 2], c3 = base[t + 3];
                mm |= within_h_k(pi, c0, k) ? (1ull << t) : 0ull;
                mm |= within_h_k(pi, c1, k) ? (2ull << t) : 0ull;
                mm |= within_h_k(pi, c2, k) ? (4ull << t) : 0ull;
                mm |= within_h_k(pi, c3, k) ? (8ull << t) : 0ull;
            }
            for (; t < L; ++t) mm |= within_h_k(pi, base[t], k) ? (1ull << t) : 0ull;
            masks[(size_t)r * n + i] = mm;
        } else {
            mm = masks[(size_t)r * n + i];
        }
        if (__any((int)((mm >> 32) != 0ull)))
            visit_row<unsigned long long>(mm, j0, delta, src, src_aux, f);
        else
            visit_row<unsigned int>((unsigned int)mm, j0, delta, src, src_aux, f);
    }
}

template <bool kBuild, class F>
__device__ __forceinline__ void for_each_neighbor_tiled(const TileShared& S, int i, int n, int* __restrict__ key2,
                                                        unsigned long long* __restrict__ masks,
                                                        const float4* __restrict__ xp,
                                                        const float4* __restrict__ cur,
                                                        const float4* __restrict__ aux,
                                                        const float4* __restrict__ lds,
                                                        const float4* __restrict__ lds_aux, const PbfK& k,
                                                        F&& f) {
    if (S.staged)
        for_each_neighbor_tiled_impl<kBuild, true>(S, i, n, key2, masks, xp, lds, lds_aux, k, f);
    else
        for_each_neighbor_tiled_impl<kBuild, false>(S, i, n, key2, masks, xp, cur, aux, k, f);
}

// Tiles of the step: chunk q = column * nzc + c covers cells
// [c * zt, min((c + 1) * zt, gZ)) of column (y, x); its particles are cut into
// ceil(count / kTileT) tiles of consecutive slots.
__global__ __launch_bounds__(kBlock) void k_tile_count(const int* __restrict__ cs, PbfK k, int zt, int nzc,
                                                       int nchunks, int* __restrict__ tcount) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nchunks) return;
    const int col = q / nzc, c = q - col * nzc;
    const int y = col / k.gX, x = col - y * k.gX;
    const int row = y * k.gXZ + x * k.gZ;
    const int z0 = c * zt;
    const int z1 = z0 + zt < k.gZ ? z0 + zt : k.gZ;
    const int cnt = cs[row + z1] - cs[row + z0];
    tcount[q] = (cnt + kTileT - 1) / kTileT;
}

__global__ __launch_bounds__(kBlock) void k_tile_write(const int* __restrict__ cs, const int* __restrict__ key2,
                                                       PbfK k, int zt, int nzc, int nchunks,
                                                       const int* __restrict__ toff, int4* __restrict__ tiles) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nchunks) return;
    const int col = q / nzc, c = q - col * nzc;
    const int y = col / k.gX, x = col - y * k.gX;
    const int row = y * k.gXZ + x * k.gZ;
    const int z0 = c * zt;
    const int z1 = z0 + zt < k.gZ ? z0 + zt : k.gZ;
    const int s = cs[row + z0], e = cs[row + z1];
    int out = toff[q];
    for (int sb = s; sb < e; sb += kTileT, ++out) {
        const int se = sb + kTileT < e ? sb + kTileT : e;
        const int ka = key2[sb] & ~kOvf, kb = key2[se - 1] & ~kOvf;
        const int za = ka - (ka / k.gZ) * k.gZ, zb = kb - (kb / k.gZ) * k.gZ;
        tiles[out] = make_int4(sb, se, row, za | (zb << 16));
    }
}

hipError_t launch_tiles(const int* cs, const int* key2, PbfK k, int zt, int* tcount, int* toff, int* scan_tiles,
                        int4* tiles, hipStream_t s) {
    if (zt < 1 || zt > kTileZMax) return hipErrorInvalidValue;
    const int nzc = (k.gZ + zt - 1) / zt;
    const int nchunks = k.gX * k.gY * nzc;
    const dim3 g((nchunks + kBlock - 1) / kBlock), b(kBlock);
    launch_kernel(k_tile_count, g, b, 0, s, cs, k, zt, nzc, nchunks, tcount);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    e = launch_exclusive_scan(tcount, toff, nchunks + 1, scan_tiles, 1, s);
    if (e != hipSuccess) return e;
    launch_kernel(k_tile_write, g, b, 0, s, cs, key2, k, zt, nzc, nchunks, toff, tiles);
    return hipGetLastError();
}

int tile_chunks(const PbfK& k, int zt) { return k.gX * k.gY * ((k.gZ + zt - 1) / zt); }

__global__ __launch_bounds__(kBlock) void k_pbf_key(const float4* __restrict__ x,
                                                     const float4* __restrict__ v, int n, PbfK k,
                                                     float dt, int* __restrict__ key,
                                                     int* __restrict__ slot,
                                                     int* __restrict__ counts) {
    const int i = xcd_index();
    const bool active = i < n;
    int c = 0;
    if (active) {
        const Bounds b = bounds_of(k, k.scenes ? k.sid[i] : 0);
        // solids of a mixed scene keep their position
        const int ph = k.ph ? k.ph[i] : 0;
        const float4 xs = ph == 2 ? x[i] : pbf_predict(x[i], v[i], k, b, dt);
        c = pbf_cell(xs, k, b.yoff, k.nph == 1 ? 0 : ph);
        key[i] = c;
    }
    const int sl = wave_atomic_slot(counts, c, active);
    if (active) slot[i] = sl;
}

hipError_t launch_pbf_key(const float4* x, const float4* v, int n, PbfK k, float dt, int* key,
                          int* slot, int* counts, hipStream_t s) {
    launch_kernel(k_pbf_key, dim3(grid_for(n)), dim3(kBlock), 0, s, x, v, n, k, dt, key, slot,
                       counts);
    return hipGetLastError();
}

__global__ __launch_bounds__(kBlock) void k_pbf_permute(
    const int* __restrict__ tmp, const int* __restrict__ key, const int* __restrict__ cell_start,
    const float4* __restrict__ x, const float4* __restrict__ v, const int* __restrict__ id, int n,
    PbfK k, float dt, float4* __restrict__ x2, float4* __restrict__ xp, int* __restrict__ id2,
    int* __restrict__ key2, const int* __restrict__ gate) {
    if (gate && *gate == 0) return;
    const int p = xcd_index();
    if (p >= n) return;
    const int i = tmp[p];
    const int c = key[i];
    const int o = stable_slot(tmp, cell_start, c, i);
    const float4 xi = x[i];
    const int s = k.scenes ? k.sid[i] : 0;
    const int ph = k.ph ? k.ph[i] : 0;
    float4 xq = ph == 2 ? xi : pbf_predict(xi, v[i], k, bounds_of(k, s), dt);
    if (k.ph) {
        // mixed scene: the phase rides in w of the old and predicted positions
        // (and of the corrected ones), so neighbour loops read it with the
        // position instead of gathering ph_out[j]
        const float fph = (float)ph;
        x2[o] = make_float4(xi.x, xi.y, xi.z, fph);
        xq.w = fph;
    } else {
        x2[o] = xi;
    }
    xp[o] = xq;
    id2[o] = id[i];
    key2[o] = c;
    if (k.scenes) k.sid_out[o] = s;
    if (k.ph) k.ph_out[o] = ph;
}

hipError_t launch_pbf_permute(const int* tmp, const int* key, const int* cell_start,
                              const float4* x, const float4* v, const int* id, int n, PbfK k,
                              float dt, float4* x2, float4* xp, int* id2, int* key2,
                              const int* gate, hipStream_t s) {
    launch_kernel(k_pbf_permute, dim3(grid_for(n)), dim3(kBlock), 0, s, tmp, key, cell_start,
                       x, v, id, n, k, dt, x2, xp, id2, key2, gate);
    return hipGetLastError();
}

// x2 = the step's old positions in cell order again (the permute's x2 output
// alone), for the plain rerun of a learned step
__global__ __launch_bounds__(kBlock) void k_pbf_restore_old(const int* __restrict__ tmp,
                                                            const int* __restrict__ key,
                                                            const int* __restrict__ cell_start,
                                                            const float4* __restrict__ x, int n,
                                                            float4* __restrict__ x2,
                                                            const int* __restrict__ gate) {
    if (gate && *gate == 0) return;
    const int p = xcd_index();
    if (p >= n) return;
    const int i = tmp[p];
    x2[stable_slot(tmp, cell_start, key[i], i)] = x[i];
}

hipError_t launch_pbf_restore_old(const int* tmp, const int* key, const int* cell_start,
                                  const float4* x, int n, float4* x2, const int* gate, hipStream_t s) {
    launch_kernel(k_pbf_restore_old, dim3(grid_for(n)), dim3(kBlock), 0, s, tmp, key, cell_start, x, n,
                       x2, gate);
    return hipGetLastError();
}

// s(r) = SPf (hf - f r)^2 / r of the spiky gradient, with r = |d|
template <bool kFast>
__device__ __forceinline__ float spiky_s(float r2, const PbfK& k) {
    if (kFast) {
        const float rinv = __builtin_amdgcn_rsqf(r2);
        const float rl = r2 * rinv;
        const float g = __builtin_fmaf(-k.f, rl, k.hf);
        return (k.SPf * g) * (g * rinv);
    } else {
        const float rl = sqrtf(r2);
        const float g = k.hf - k.f * rl;
        return ((k.SPf * g) * g) / rl;
    }
}

template <bool kFast>
__device__ __forceinline__ float dist2(float dx, float dy, float dz) {
    if (kFast) return __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
    return dx * dx + dy * dy + dz * dz;
}

template <bool kFast>
__device__ __forceinline__ float acc(float sum, float a, float b) {  // sum + a * b
    if (kFast) return __builtin_fmaf(a, b, sum);
    return sum + a * b;
}

template <bool kFast>
__device__ __forceinline__ float poly_t(float r2, const PbfK& k) {  // hf^2 - f^2 r^2
    if (kFast) return __builtin_fmaf(-k.f2, r2, k.hf2);
    return k.hf2 - k.f2 * r2;
}

// Block max of C, one plain store per block (out[blockIdx.x]); the consumer
// reduces the few thousand block values when it needs them.  Every thread of
// the block must call it.
__device__ __forceinline__ void block_max_store(float c, int* __restrict__ out) {
    __shared__ int wmax[kBlock / 64];
    int m = ordered(c);
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        int o = __shfl_xor(m, d, 64);
        m = o > m ? o : m;
    }
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        int b = wmax[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) b = wmax[w] > b ? wmax[w] : b;
        out[blockIdx.x] = b;
    }
}

// block_max_store of cmax plus the block sum of the threads' csum (fixed
// reduction tree), as float bits at out[gridDim.x + blockIdx.x]
__device__ __forceinline__ void block_max_sum_store(float cmax, float csum, int* __restrict__ out) {
    __shared__ float wsum[kBlock / 64];
    float s = csum;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) s += __shfl_xor(s, d, 64);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = s;
    block_max_store(cmax, out);  // contains the barrier that publishes wsum
    if (threadIdx.x == 0) {
        float b = 0.0f;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) b += wsum[w];
        out[gridDim.x + blockIdx.x] = __float_as_int(b);
    }
}

// Each neighbour pass is a per-particle body run by two kernels: one thread
// per slot (flat) or one tile per block with LDS staging (tiled, suffix _t).
// `visit(f)` calls f(j, cur_j, aux_j) over N(i) in ascending slot order.

// A uniform kernel constant held in a vector register for the neighbour loop
// (an instruction takes one scalar operand: a second one would be copied to
// a vector register on every iteration).
__device__ __forceinline__ float vreg(float x) {
    float v;
    asm volatile("v_mov_b32 %0, %1" : "=v"(v) : "s"(x));
    return v;
}

// Fused-mode spiky factor without SPf: g^2 / r with g = hf - f r,
// from rinv = 1 / sqrt(r2) (v_rsq_f32)
__device__ __forceinline__ float spiky_unit(float r2, float mf, float hf) {
    const float rinv = __builtin_amdgcn_rsqf(r2);
    const float g = __builtin_fmaf(mf, r2 * rinv, hf);
    return g * (g * rinv);
}

// lambda pass: density constraint and its gradient norm (PBF eq. 1, 8, 11),
// summed over the neighbours pj of pi (add), then C_i and lambda_i (finish)
template <bool kFast, bool kVreg = true>
struct LambdaAcc {
    float rho = 0.0f, a = 0.0f, gx = 0.0f, gy = 0.0f, gz = 0.0f;
    float mf, hf, mf2, hf2;  // fast mode: kernel constants (kVreg: in vector registers)

    __device__ __forceinline__ explicit LambdaAcc(const PbfK& k) {
        if constexpr (kFast && kVreg) {
            mf = vreg(-k.f);
            hf = vreg(k.hf);
            mf2 = vreg(-k.f2);
            hf2 = vreg(k.hf2);
        } else if constexpr (kFast) {
            mf = -k.f;
            hf = k.hf;
            mf2 = -k.f2;
            hf2 = k.hf2;
        }
    }
    __device__ __forceinline__ void add(const float4 pi, const float4 pj, const PbfK& k) {
        const float dx = pi.x - pj.x;
        const float dy = pi.y - pj.y;
        const float dz = pi.z - pj.z;
        if constexpr (kFast) {
            // SPf leaves the sums: a and |grad|^2 scale by SPf^2 (f_mr0sq_spf2)
            const float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
            if (r2 <= k.h2) {
                const float t = __builtin_fmaf(mf2, r2, hf2);
                rho = __builtin_fmaf(t * t, t, rho);
                if (r2 > 0.0f) {
                    const float s = spiky_unit(r2, mf, hf);
                    a = __builtin_fmaf(s * s, r2, a);
                    gx = __builtin_fmaf(s, dx, gx);
                    gy = __builtin_fmaf(s, dy, gy);
                    gz = __builtin_fmaf(s, dz, gz);
                }
            }
        } else {
            const float r2 = dist2<kFast>(dx, dy, dz);
            if (r2 <= k.h2) {
                const float t = poly_t<kFast>(r2, k);
                rho = acc<kFast>(rho, t * t, t);
                if (r2 > 0.0f) {
                    const float s = spiky_s<kFast>(r2, k);
                    a = acc<kFast>(a, s * s, r2);
                    gx = acc<kFast>(gx, s, dx);
                    gy = acc<kFast>(gy, s, dy);
                    gz = acc<kFast>(gz, s, dz);
                }
            }
        }
    }
    // C_i; writes (x_i, lambda_i) to *out
    __device__ __forceinline__ float finish(const float4 pi, float4* __restrict__ out, const PbfK& k) const {
        const float Ci = rho * k.c_rho - 1.0f;
        const float gg = gx * gx + gy * gy + gz * gz;
        const float den = (kFast ? k.f_mr0sq_spf2 : k.mr0sq) * (a + gg) + k.eps;
        *out = make_float4(pi.x, pi.y, pi.z, -Ci / den);
        return Ci;
    }
};

template <bool kFast, class Visit>
__device__ __forceinline__ float lambda_body(int i, const float4* __restrict__ xa, float4* __restrict__ xl,
                                             const PbfK& k, Visit&& visit) {
    const float4 pi = xa[i];
    LambdaAcc<kFast> s(k);
    visit([&](int, const float4 pj, const float4) { s.add(pi, pj, k); });
    return s.finish(pi, xl + i, k);
}

template <bool kBuild, bool kFast>
__global__ __launch_bounds__(kBlock) void k_pbf_lambda(const float4* __restrict__ xa,
                                                        int* __restrict__ key2,
                                                        const int* __restrict__ cs, int n, PbfK k,
                                                        float4* __restrict__ xl,
                                                        int* __restrict__ maxc,
                                                        unsigned long long* __restrict__ masks,
                                                        const float4* __restrict__ xp,
                                                        const int* __restrict__ gate) {
    if (gate && *gate == 0) return;  // uniform: before any barrier
    const int i = xcd_index();
    float Ci = -1e30f;
    if (i < n)
        Ci = lambda_body<kFast>(i, xa, xl, k, [&](auto&& f) {
            for_each_neighbor<kBuild, kWalkWide>(i, n, key2, cs, masks, xp, xa, nullptr, k, f);
        });
    block_max_store(Ci, maxc);
}

// Wave-staged list build: the first lambda pass of a step, testing the
// candidates at x* (the pass's own input).  The 64 slots of a wave lie in a
// few (y, x) cell columns; within a column the slots ascend with cz and so do
// their stencil windows, so the union of the wave's windows in stencil row r
// is one contiguous slot range per column.  The wave loads those ranges once,
// coalesced, into LDS and its lanes test their candidates and read
// their hits there instead of gathering them one lane at a time through the
// texture units.  The staged copy holds x, y, z of a slot in three consecutive
// floats, so one address reaches four candidates through the instructions'
// offset fields.  A row whose ranges exceed kStageCap slots or kStageSegs
// columns is tested from global memory as build_list does.  Same candidates,
// same order, same arithmetic as build_list + lambda_body: the same bits.
constexpr int kStageCap = 256;
constexpr int kStageSegs = 4;

// staged slot q (kS floats: x, y, z and, for kS = 4, w)
template <int kS>
__device__ __forceinline__ float4 staged_at(const lds_float* __restrict__ S, int q) {
    const lds_float* __restrict__ p = S + __umul24((unsigned)q, (unsigned)kS);  // q < kStageCap
    return make_float4(p[0], p[1], p[2], kS == 4 ? p[3] : 0.0f);
}

// Visit the set bits of a row mask in ascending order from the staged copy:
// f(j0 + t, S[lo + t]) (LDS latency is short: no read ahead)
template <typename M, int kS, class F>
__device__ __forceinline__ void visit_row_staged(M mm, int j0, int lo, const lds_float* __restrict__ S, F& f) {
    while (mm) {
        const int t = low_bit(mm);
        mm &= mm - (M)1;
        f(j0 + t, staged_at<kS>(S, lo + t));
    }
}

__device__ __forceinline__ int high_bit(unsigned int m) { return 31 - __builtin_clz(m); }
__device__ __forceinline__ int high_bit(unsigned long long m) { return 63 - __builtin_clzll(m); }

// o shifted left by one with bit 0 = "staged candidate p is outside h of pi":
// the sign of h2 - r2 as within_h<kFast> forms it.  Exact mode: the rounded
// difference of two floats has the sign of their exact difference and is +0
// when they are equal, so the bit is 1 exactly when r2 > h2.  Fast mode: the
// fused multiply-adds give -0 never (an exact zero sum rounds to +0), so the
// bit is 1 exactly when h2 - r2 < 0.  One funnel shift replaces the compare,
// the select and the or of a mask update.
template <int kS, bool kFast>
__device__ __forceinline__ unsigned int shift_outside(unsigned int o, const float4 pi, const lds_float* __restrict__ p,
                                                      const PbfK& k) {
    const float dx = pi.x - p[0];
    const float dy = pi.y - p[1];
    const float dz = pi.z - p[2];
    if constexpr (kFast) return __builtin_amdgcn_alignbit(o, __float_as_uint(h2_less_r2_fast(dx, dy, dz, k)), 31u);
    const float r2 = dx * dx + dy * dy + dz * dz;
    return __builtin_amdgcn_alignbit(o, __float_as_uint(k.h2 - r2), 31u);
}

// Hit mask of the cnt (<= 32) staged candidates b[t]: shifted in from the
// last to the first, so candidate t lands in bit t; the all-ones start leaves
// ones (no hit) above bit cnt - 1.
template <int kS, bool kFast>
__device__ __forceinline__ unsigned int hits32_staged(const float4 pi, const lds_float* __restrict__ b, int cnt,
                                                      const PbfK& k) {
    unsigned int o = ~0u;
    int t = cnt;
    for (; t >= 4; t -= 4) {
        const lds_float* __restrict__ q = b + kS * (t - 4);
        o = shift_outside<kS, kFast>(o, pi, q + 3 * kS, k);
        o = shift_outside<kS, kFast>(o, pi, q + 2 * kS, k);
        o = shift_outside<kS, kFast>(o, pi, q + kS, k);
        o = shift_outside<kS, kFast>(o, pi, q, k);
    }
    for (; t > 0; --t) o = shift_outside<kS, kFast>(o, pi, b + kS * (t - 1), k);
    return ~o;
}

// Hit mask of the L (<= bits of M) staged candidates S[lo + t] of pi
template <typename M, int kS, bool kFast>
__device__ __forceinline__ M test_row_staged(const float4 pi, int lo, int L, const lds_float* __restrict__ S,
                                             const PbfK& k) {
    const lds_float* __restrict__ b = S + kS * lo;
    if constexpr (sizeof(M) == 4) {
        return hits32_staged<kS, kFast>(pi, b, L, k);
    } else {
        const int nl = L < 32 ? L : 32;
        const unsigned int l = hits32_staged<kS, kFast>(pi, b, nl, k);
        const unsigned int h = hits32_staged<kS, kFast>(pi, b + 32 * kS, L - nl, k);
        return (M)h << 32 | (M)l;
    }
}

// The mixed scene's lambda sums (fluid i): rho and the gradient over every
// neighbour, the own-gradient-norm term over fluid neighbours (w == 0) only
template <bool kFast>
struct MixLambdaAcc {
    float rho = 0.0f, a = 0.0f, gx = 0.0f, gy = 0.0f, gz = 0.0f;
    __device__ __forceinline__ void add(const float4 pi, const float4 pj, const PbfK& k) {
        const float dx = pi.x - pj.x;
        const float dy = pi.y - pj.y;
        const float dz = pi.z - pj.z;
        const float r2 = dist2<kFast>(dx, dy, dz);
        if (r2 <= k.h2) {
            const float t = poly_t<kFast>(r2, k);
            rho = acc<kFast>(rho, t * t, t);
            if (r2 > 0.0f) {
                const float s = spiky_s<kFast>(r2, k);
                if (pj.w == 0.0f) a = acc<kFast>(a, s * s, r2);
                gx = acc<kFast>(gx, s, dx);
                gy = acc<kFast>(gy, s, dy);
                gz = acc<kFast>(gz, s, dz);
            }
        }
    }
    __device__ __forceinline__ float lambda(const PbfK& k) const {
        const float Ci = rho * k.c_rho - 1.0f;
        const float gg = gx * gx + gy * gy + gz * gz;
        const float den = k.mr0sq * (a + gg) + k.eps;
        return -Ci / den;
    }
};

// kMixed: the mixed scene's first lambda pass (phase runs, nph = 3): fluid
// lanes sum MixLambdaAcc, sand lanes only build their lists, solid lanes
// write an empty list (no pass reads it); the staged copy keeps w (the phase).
template <bool kFast, bool kMixed>
__device__ __forceinline__ void lambda_build_ws(const float4* __restrict__ xp, const int* __restrict__ key2,
                                                const int* __restrict__ cs, int n, const PbfK& k,
                                                float4* __restrict__ xl, int* __restrict__ maxc,
                                                unsigned long long* __restrict__ nb, const int* __restrict__ gate) {
    constexpr int kS = kMixed ? 4 : 3;  // floats per staged slot
    // runs per wave staged (a mixed wave holds a column's fluid, sand and
    // solid runs, so more of them)
    constexpr int kSegs = kMixed ? 2 * kStageSegs : kStageSegs;
    if (gate && *gate == 0) return;  // uniform: before any barrier
    __shared__ float stage[kBlock / 64][kS * kStageCap];
    const int i = xcd_index();
    const int lane = (int)(threadIdx.x & 63);
    lds_float* __restrict__ S = (lds_float*)stage[threadIdx.x >> 6];
    // every lane of the wave takes part in the staging; a lane past n
    // (i >= n, the last wave only) has empty windows and stores nothing
    const bool act = i < n;
    const int ii = act ? i : n - 1;
    const int ph = kMixed ? k.ph_out[ii] : 0;
    const bool lists = act && ph != 2;
    int cy, cx, cz;
    key_cell(key2[ii] & ~kOvf, k, cy, cx, cz);
    const int z0 = cz > 0 ? cz - 1 : 0;
    const int z1 = cz + 1 < k.gZ ? cz + 1 : k.gZ - 1;
    const float4 pi = xp[ii];
    int4* __restrict__ lst = nbr_list(nb, n, ii);
    LambdaAcc<kFast, false> acc(k);
    MixLambdaAcc<kFast> macc;
    ListWriter lw(ii);
    auto sum = [&](const float4 pj) {
        if constexpr (kMixed) {
            if (ph == 0) macc.add(pi, pj, k);
        } else {
            acc.add(pi, pj, k);
        }
    };
    auto hit = [&](int j, const float4 pj) {
        lw.emit(j, lst);
        sum(pj);
    };
    // a hit of a row announced with row_hits
    auto hit_row = [&](int j, const float4 pj) {
        if (lw.append_row(j)) lst[(size_t)((lw.c >> 3) - 1) * 64] = lw.group();
        sum(pj);
    };
    const unsigned long long below = (1ull << lane) - 1ull;  // lanes before this one
#pragma unroll 1
    for (int r = 0; r < 9; ++r) {
        lw.row(r, ii);
        const int y = cy + r / 3 - 1;
        const int x = cx + r % 3 - 1;
        const bool valid = lists && y >= 0 && y < k.gY && x >= 0 && x < k.gX;
        // runs of lanes whose row r lies in one column and whose own phase
        // run is one (the slots of a run ascend with cz, and so do its
        // windows); -1: no row
        const int col = valid ? (y * k.gX + x) * k.nph + ph : -1;
        const int pcol = __shfl_up(col, 1, 64);
        const int ncol = __shfl_down(col, 1, 64);
        const unsigned long long fm = __ballot(valid && (lane == 0 || pcol != col));
        const unsigned long long lm = __ballot(valid && (lane == 63 || ncol != col));
        const int nseg = __popcll(fm);
        if (nseg == 0) continue;  // uniform: no lane has row r
#pragma unroll 1
        for (int c = 0; c < k.nph; ++c) {  // the row's phase runs (one range for nph = 1)
            int j0 = 0, j1 = 0;
            if (valid) {
                const int row = row_key(y, x, c, k);
                j0 = cs[row + z0];
                j1 = cs[row + z1 + 1];
            }
            const int L = j1 - j0;
            // the runs' slot ranges [start_s, end_s), packed at LDS offsets off_s
            int s_start[kSegs], s_off[kSegs + 1];
            int tot = 0;
            {
                unsigned long long f = fm, l = lm;
#pragma unroll
                for (int s = 0; s < kSegs; ++s) {
                    s_off[s] = tot;
                    s_start[s] = 0;
                    if (s < nseg) {
                        const int lf = __ffsll((long long)f) - 1, ll = __ffsll((long long)l) - 1;
                        f &= f - 1ull;
                        l &= l - 1ull;
                        const int a = __builtin_amdgcn_readlane(j0, lf);
                        const int b = __builtin_amdgcn_readlane(j1, ll);
                        s_start[s] = a;
                        tot += b - a;
                    }
                }
                s_off[kSegs] = tot;
            }
            // uniform: every window of this range is empty (tot covers every
            // run only when there are at most kSegs of them)
            if (tot == 0 && nseg <= kSegs) continue;
            const bool staged = nseg <= kSegs && tot <= kStageCap;  // uniform
            if (staged) {
                for (int q = lane; q < tot; q += 64) {
                    int j = q - s_off[0] + s_start[0];
                    // uniform: one run needs no selection, two runs (the mixed
                    // scene's usual fluid and sand runs) one compare; the plain
                    // build keeps the general form (its registers: 80, 6 waves)
                    if (kMixed && nseg == 2) {
                        if (q >= s_off[1]) j = q - s_off[1] + s_start[1];
                    } else if (nseg > 1) {
#pragma unroll
                        for (int s = 1; s < kSegs; ++s)
                            if (q >= s_off[s]) j = q - s_off[s] + s_start[s];
                    }
                    const float4 p = at4(xp, j);
                    S[kS * q] = p.x;
                    S[kS * q + 1] = p.y;
                    S[kS * q + 2] = p.z;
                    if (kS == 4) S[kS * q + 3] = p.w;
                }
                wave_lds_sync();
                // this lane's run and the LDS offset of its window
                int lo = j0 - s_start[0] + s_off[0];
                if (nseg > 1) {  // uniform
                    const int seg = __popcll(fm & (below | (1ull << lane))) - 1;
#pragma unroll
                    for (int s = 1; s < kSegs; ++s)
                        if (seg == s) lo = j0 - s_start[s] + s_off[s];
                }
                if (L > 64) {
                    for (int t = 0; t < L; ++t) {
                        const float4 q = staged_at<kS>(S, lo + t);
                        if (within_h<kFast>(pi, q, k)) hit(j0 + t, q);
                    }
                } else if (__all(L <= 32)) {
                    const unsigned int mm = test_row_staged<unsigned int, kS, kFast>(pi, lo, L, S, k);
                    if (mm) lw.row_hits(j0 + low_bit(mm), j0 + high_bit(mm));
                    visit_row_staged<unsigned int, kS>(mm, j0, lo, S, hit_row);
                } else {
                    const unsigned long long mm = test_row_staged<unsigned long long, kS, kFast>(pi, lo, L, S, k);
                    if (mm) lw.row_hits(j0 + low_bit(mm), j0 + high_bit(mm));
                    if (__any((int)((mm >> 32) != 0ull)))
                        visit_row_staged<unsigned long long, kS>(mm, j0, lo, S, hit_row);
                    else
                        visit_row_staged<unsigned int, kS>((unsigned int)mm, j0, lo, S, hit_row);
                }
                wave_lds_sync();  // the next range's staging overwrites this one's
            } else if (valid) {
                if (L > 64) {
                    for (int j = j0; j < j1; ++j) {
                        const float4 q = at4(xp, j);
                        if (within_h<kFast>(pi, q, k)) hit(j, q);
                    }
                    continue;
                }
                unsigned long long mm = 0ull;
                const float4* __restrict__ base = xp + j0;
                int t = 0;
                for (; t + 4 <= L; t += 4) {
                    const float3 c0 = xyz_at(base, t), c1 = xyz_at(base, t + 1), c2 = xyz_at(base, t + 2),
                                 c3 = xyz_at(base, t + 3);
                    const unsigned int nib = (within_h<kFast>(pi, c0, k) ? 1u : 0u) | (within_h<kFast>(pi, c1, k) ? 2u : 0u) |
                                             (within_h<kFast>(pi, c2, k) ? 4u : 0u) | (within_h<kFast>(pi, c3, k) ? 8u : 0u);
                    mm |= (unsigned long long)nib << t;
                }
                for (; t < L; ++t) mm |= within_h<kFast>(pi, xyz_at(base, t), k) ? (1ull << t) : 0ull;
                auto hit4 = [&](int j, const float4 pj, const float4) { hit(j, pj); };
                if (__any((int)((mm >> 32) != 0ull)))
                    visit_row<unsigned long long>(mm, j0, 0, xp, nullptr, hit4);
                else
                    visit_row<unsigned int>((unsigned int)mm, j0, 0, xp, nullptr, hit4);
            }
        }
    }
    float Ci = -1e30f;
    if (act) {
        if constexpr (kMixed) {
            xl[i] = make_float4(pi.x, pi.y, pi.z, ph == 0 ? macc.lambda(k) : 0.0f);
            if (lists)
                lw.finish(i, nbr_hdr(nb), lst);
            else
                nbr_hdr(nb)[i] = make_int4(i, i, i, 0);
        } else {
            Ci = acc.finish(pi, xl + i, k);
            lw.finish(i, nbr_hdr(nb), lst);
        }
    }
    if (!kMixed) block_max_store(Ci, maxc);
}

// the plain scene's first lambda pass (kMixed is false)
template <bool kFast, bool kMixed>
__global__ __launch_bounds__(kBlock) void k_pbf_lambda_build_ws(const float4* __restrict__ xp,
                                                                 const int* __restrict__ key2,
                                                                 const int* __restrict__ cs, int n, PbfK k,
                                                                 float4* __restrict__ xl,
                                                                 int* __restrict__ maxc,
                                                                 unsigned long long* __restrict__ nb,
                                                                 const int* __restrict__ gate) {
    lambda_build_ws<kFast, kMixed>(xp, key2, cs, n, k, xl, maxc, nb, gate);
}

// The mixed scene's: held to 72 VGPRs (7 waves per SIMD, a few registers
// spilled), measured faster there (125.7 -> 117.1 us at 400 K particles) and
// slower for the plain scene's build, which keeps 80 (6 waves).
template <bool kFast>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(7, 8))) void k_mix_lambda_build_ws(
    const float4* __restrict__ xp, const int* __restrict__ key2, const int* __restrict__ cs, int n, PbfK k,
    float4* __restrict__ xl, int* __restrict__ maxc, unsigned long long* __restrict__ nb,
    const int* __restrict__ gate) {
    lambda_build_ws<kFast, true>(xp, key2, cs, n, k, xl, maxc, nb, gate);
}

template <bool kBuild, bool kFast>
__global__ __launch_bounds__(kTileT) void k_pbf_lambda_t(const float4* __restrict__ xa,
                                                          int* __restrict__ key2,
                                                          const int* __restrict__ cs, int n, PbfK k,
                                                          float4* __restrict__ xl,
                                                          int* __restrict__ maxc,
                                                          unsigned long long* __restrict__ masks,
                                                          const float4* __restrict__ xp,
                                                          const int* __restrict__ gate,
                                                          const int4* __restrict__ tiles,
                                                          const int* __restrict__ ntiles, int lcap) {
    if (gate && *gate == 0) return;
    extern __shared__ float4 lds[];
    __shared__ TileShared S;
    float cmax = -1e30f;
    const int nt = *ntiles;
    for (int t = tile_first(); t < nt; t += gridDim.x) {
        tile_stage(S, tiles[t], cs, k, xa, nullptr, lds, nullptr, lcap);
        const int i = S.i0 + (int)threadIdx.x;
        if (i < S.i1) {
            const float c = lambda_body<kFast>(i, xa, xl, k, [&](auto&& f) {
                for_each_neighbor_tiled<kBuild>(S, i, n, key2, masks, xp, xa, nullptr, lds, nullptr, k, f);
            });
            cmax = c > cmax ? c : cmax;
        }
        __syncthreads();  // the next tile overwrites the staged data
    }
    block_max_store(cmax, maxc);
}

hipError_t launch_pbf_lambda(const float4* xa, int* key2, const int* cs, int n, PbfK k, float4* xl,
                             int* maxc, unsigned long long* masks, const float4* xp, int build,
                             const int* gate, const TileArgs& ta, hipStream_t s) {
    // a building pass tests candidates against the staged input: it must be x*
    if (ta.tiles && (!build || xa == xp)) {
        const dim3 g(ta.grid), b(kTileT);
        const size_t sh = (size_t)ta.lcap * sizeof(float4);
        if (build && k.exact)
            launch_kernel((k_pbf_lambda_t<true, false>), g, b, sh, s, xa, key2, cs, n, k, xl, maxc, masks, xp,
                               gate, ta.tiles, ta.ntiles, ta.lcap);
        else if (build)
            launch_kernel((k_pbf_lambda_t<true, true>), g, b, sh, s, xa, key2, cs, n, k, xl, maxc, masks, xp,
                               gate, ta.tiles, ta.ntiles, ta.lcap);
        else if (k.exact)
            launch_kernel((k_pbf_lambda_t<false, false>), g, b, sh, s, xa, key2, cs, n, k, xl, maxc, masks, xp,
                               gate, ta.tiles, ta.ntiles, ta.lcap);
        else
            launch_kernel((k_pbf_lambda_t<false, true>), g, b, sh, s, xa, key2, cs, n, k, xl, maxc, masks, xp,
                               gate, ta.tiles, ta.ntiles, ta.lcap);
        return hipGetLastError();
    }
    if (!build && ta.sub > 1 && !k.exact) return launch_lambda_sub(xa, key2, cs, n, k, xl, maxc, masks, xp, gate, s);
    if (build && ta.sub > 1 && !k.exact && xa == xp && k.nph == 1 && n > 0)
        return launch_lambda_build_sub(xp, key2, cs, n, k, xl, maxc, masks, gate, s);
    const dim3 g(grid_for(n, ta.block)), b(ta.block);
    if (build && ta.stage && xa == xp && n > 0 && k.nph == 1) {
        if (k.exact)
            launch_kernel((k_pbf_lambda_build_ws<false, false>), g, b, 0, s, xp, key2, cs, n, k, xl, maxc, masks,
                          gate);
        else
            launch_kernel((k_pbf_lambda_build_ws<true, false>), g, b, 0, s, xp, key2, cs, n, k, xl, maxc, masks,
                          gate);
        return hipGetLastError();
    }
    if (build && k.exact)
        launch_kernel((k_pbf_lambda<true, false>), g, b, 0, s, xa, key2, cs, n, k, xl, maxc, masks, xp, gate);
    else if (build)
        launch_kernel((k_pbf_lambda<true, true>), g, b, 0, s, xa, key2, cs, n, k, xl, maxc, masks, xp, gate);
    else if (k.exact)
        launch_kernel((k_pbf_lambda<false, false>), g, b, 0, s, xa, key2, cs, n, k, xl, maxc, masks, xp, gate);
    else
        launch_kernel((k_pbf_lambda<false, true>), g, b, 0, s, xa, key2, cs, n, k, xl, maxc, masks, xp, gate);
    return hipGetLastError();
}

// delta-p sums with s_corr (PBF eq. 12-14) over the neighbours pj of pi
// (lambda_i li): add; the position step is (ax, ay, az) * step.
// kNS: s_corr exponent fixed at compile time (4), or 0 for k.n_scorr.
template <bool kFast, int kNS>
struct DeltapAcc {
    float ax = 0.0f, ay = 0.0f, az = 0.0f;
    float step;
    float mf, hf, mf2, hf2;  // fast mode, kNS = 4: kernel constants in vector registers

    __device__ __forceinline__ explicit DeltapAcc(const PbfK& k) {
        if constexpr (kFast && kNS == 4) {
            // SPf leaves the sums (f_spf_neg_inv_r0); s_corr = -k (t / t(dq))^12
            // as one constant times t^12 (f_scorr12)
            mf = vreg(-k.f);
            hf = vreg(k.hf);
            mf2 = vreg(-k.f2);
            hf2 = vreg(k.hf2);
            step = k.f_spf_neg_inv_r0;
        } else {
            step = k.neg_inv_r0;
        }
    }
    __device__ __forceinline__ void add(const float4 pi, float li, const float4 pj, const PbfK& k) {
        const float dx = pi.x - pj.x;
        const float dy = pi.y - pj.y;
        const float dz = pi.z - pj.z;
        if constexpr (kFast && kNS == 4) {
            const float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
            if (r2 <= k.h2 && r2 > 0.0f) {
                const float s = spiky_unit(r2, mf, hf);
                const float t = __builtin_fmaf(mf2, r2, hf2);
                const float t2 = t * t;
                const float t4 = t2 * t2;
                const float t12 = (t4 * t4) * t4;
                const float c = __builtin_fmaf(k.f_scorr12, t12, li + pj.w) * s;
                ax = __builtin_fmaf(c, dx, ax);
                ay = __builtin_fmaf(c, dy, ay);
                az = __builtin_fmaf(c, dz, az);
            }
        } else {
            const float r2 = dist2<kFast>(dx, dy, dz);
            if (r2 <= k.h2 && r2 > 0.0f) {
                const float s = spiky_s<kFast>(r2, k);
                const float t = poly_t<kFast>(r2, k);
                const float ratio = ((t * t) * t) * k.inv_tq3;
                float pw;
                if (kNS == 4) {
                    pw = ratio * ratio;
                    pw = pw * ratio;
                    pw = pw * ratio;
                } else {
                    pw = ratio;
                    for (int e = 1; e < k.n_scorr; ++e) pw = pw * ratio;
                }
                const float sc = -k.scorr_k * pw;
                const float coef = (li + pj.w) + sc;
                const float c = coef * s;
                ax = acc<kFast>(ax, c, dx);
                ay = acc<kFast>(ay, c, dy);
                az = acc<kFast>(az, c, dz);
            }
        }
    }
};

// x* = clamp(x* + (ax, ay, az) step) and, when vout is set, v = (x* - x) / dt
__device__ __forceinline__ void deltap_commit(int i, const float4 pi, float ax, float ay, float az, float step,
                                              const PbfK& k, float4* __restrict__ xa,
                                              const float4* __restrict__ xold, float4* __restrict__ vout,
                                              float dt) {
    const Bounds b = bounds_of(k, k.scenes ? k.sid_out[i] : 0);
    float4 r;
    r.x = rclamp(pi.x + ax * step, b.lo[0], b.hi[0]);
    r.y = rclamp(pi.y + ay * step, b.lo[1], b.hi[1]);
    r.z = rclamp(pi.z + az * step, b.lo[2], b.hi[2]);
    r.w = 0.0f;
    xa[i] = r;
    if (vout) {
        const float4 xo = xold[i];
        vout[i] = make_float4((r.x - xo.x) / dt, (r.y - xo.y) / dt, (r.z - xo.z) / dt, 0.0f);
    }
}

// delta-p pass: the sums of DeltapAcc, Jacobi update, box clamp
template <bool kFast, int kNS, class Visit>
__device__ __forceinline__ void deltap_body(int i, const float4* __restrict__ xl, const PbfK& k,
                                            float4* __restrict__ xa, const float4* __restrict__ xold,
                                            float4* __restrict__ vout, float dt, Visit&& visit) {
    const float4 pi = xl[i];
    const float li = pi.w;
    DeltapAcc<kFast, kNS> d(k);
    visit([&](int, const float4 pj, const float4) { d.add(pi, li, pj, k); });
    deltap_commit(i, pi, d.ax, d.ay, d.az, d.step, k, xa, xold, vout, dt);
}

template <bool kFast, int kNS>
__global__ __launch_bounds__(kBlock) void k_pbf_deltap(const float4* __restrict__ xl,
                                                        int* __restrict__ key2,
                                                        const int* __restrict__ cs, int n, PbfK k,
                                                        float4* __restrict__ xa,
                                                        const float4* __restrict__ xold,
                                                        float4* __restrict__ vout, float dt,
                                                        unsigned long long* __restrict__ masks,
                                                        const float4* __restrict__ xp,
                                                        const int* __restrict__ gate) {
    if (gate && *gate == 0) return;
    const int i = xcd_index();
    if (i >= n) return;
    deltap_body<kFast, kNS>(i, xl, k, xa, xold, vout, dt, [&](auto&& f) {
        for_each_neighbor<false, kWalkWide>(i, n, key2, cs, masks, xp, xl, nullptr, k, f);
    });
}

template <bool kFast, int kNS>
__global__ __launch_bounds__(kTileT) void k_pbf_deltap_t(const float4* __restrict__ xl,
                                                          int* __restrict__ key2,
                                                          const int* __restrict__ cs, int n, PbfK k,
                                                          float4* __restrict__ xa,
                                                          const float4* __restrict__ xold,
                                                          float4* __restrict__ vout, float dt,
                                                          unsigned long long* __restrict__ masks,
                                                          const float4* __restrict__ xp,
                                                          const int* __restrict__ gate,
                                                          const int4* __restrict__ tiles,
                                                          const int* __restrict__ ntiles, int lcap) {
    if (gate && *gate == 0) return;
    extern __shared__ float4 lds[];
    __shared__ TileShared S;
    const int nt = *ntiles;
    for (int t = tile_first(); t < nt; t += gridDim.x) {
        tile_stage(S, tiles[t], cs, k, xl, nullptr, lds, nullptr, lcap);
        const int i = S.i0 + (int)threadIdx.x;
        if (i < S.i1)
            deltap_body<kFast, kNS>(i, xl, k, xa, xold, vout, dt, [&](auto&& f) {
                for_each_neighbor_tiled<false>(S, i, n, key2, masks, xp, xl, nullptr, lds, nullptr, k, f);
            });
        __syncthreads();
    }
}

hipError_t launch_pbf_deltap(const float4* xl, int* key2, const int* cs, int n, PbfK k, float4* xa,
                             const float4* xold, float4* vout, float dt, unsigned long long* masks,
                             const float4* xp, const int* gate, const TileArgs& ta, hipStream_t s) {
    const bool n4 = k.n_scorr == 4;
    if (ta.tiles) {
        const dim3 g(ta.grid), b(kTileT);
        const size_t sh = (size_t)ta.lcap * sizeof(float4);
        if (k.exact && n4)
            launch_kernel((k_pbf_deltap_t<false, 4>), g, b, sh, s, xl, key2, cs, n, k, xa, xold, vout, dt, masks,
                               xp, gate, ta.tiles, ta.ntiles, ta.lcap);
        else if (k.exact)
            launch_kernel((k_pbf_deltap_t<false, 0>), g, b, sh, s, xl, key2, cs, n, k, xa, xold, vout, dt, masks,
                               xp, gate, ta.tiles, ta.ntiles, ta.lcap);
        else if (n4)
            launch_kernel((k_pbf_deltap_t<true, 4>), g, b, sh, s, xl, key2, cs, n, k, xa, xold, vout, dt, masks,
                               xp, gate, ta.tiles, ta.ntiles, ta.lcap);
        else
            launch_kernel((k_pbf_deltap_t<true, 0>), g, b, sh, s, xl, key2, cs, n, k, xa, xold, vout, dt, masks,
                               xp, gate, ta.tiles, ta.ntiles, ta.lcap);
        return hipGetLastError();
    }
    if (ta.sub > 1 && !k.exact) return launch_deltap_sub(xl, key2, cs, n, k, xa, xold, vout, dt, masks, xp, gate, s);
    const dim3 g(grid_for(n, ta.block)), b(ta.block);
    if (k.exact && n4)
        launch_kernel((k_pbf_deltap<false, 4>), g, b, 0, s, xl, key2, cs, n, k, xa, xold, vout, dt, masks, xp, gate);
    else if (k.exact)
        launch_kernel((k_pbf_deltap<false, 0>), g, b, 0, s, xl, key2, cs, n, k, xa, xold, vout, dt, masks, xp, gate);
    else if (n4)
        launch_kernel((k_pbf_deltap<true, 4>), g, b, 0, s, xl, key2, cs, n, k, xa, xold, vout, dt, masks, xp, gate);
    else
        launch_kernel((k_pbf_deltap<true, 0>), g, b, 0, s, xl, key2, cs, n, k, xa, xold, vout, dt, masks, xp, gate);
    return hipGetLastError();
}

__global__ __launch_bounds__(kBlock) void k_pbf_velocity(const float4* __restrict__ xa,
                                                          const float4* __restrict__ xold, int n,
                                                          float dt, float4* __restrict__ vout) {
    const int i = xcd_index();
    if (i >= n) return;
    const float4 r = xa[i];
    const float4 xo = xold[i];
    vout[i] = make_float4((r.x - xo.x) / dt, (r.y - xo.y) / dt, (r.z - xo.z) / dt, 0.0f);
}

hipError_t launch_pbf_velocity(const float4* xa, const float4* xold, int n, float dt,
                               float4* vout, hipStream_t s) {
    launch_kernel(k_pbf_velocity, dim3(grid_for(n)), dim3(kBlock), 0, s, xa, xold, n, dt, vout);
    return hipGetLastError();
}

// XSPH viscosity (PBF eq. 17) over N(i) and commit x = x*.  kCheck: the same
// kernel sums give the density at the final positions; its constraint's block
// maximum goes to maxd (the learned step's fallback test and the density
// error statistics).
template <bool kFast, bool kCheck, class Visit>
__device__ __forceinline__ float xsph_body(int i, const float4* __restrict__ xa, const float4* __restrict__ vt,
                                           const PbfK& k, float4* __restrict__ vout, float4* __restrict__ xout,
                                           Visit&& visit, float4* xn = nullptr, float4* vn = nullptr) {
    const float4 pi = xa[i];
    const float4 vi = vt[i];
    float sx = 0.0f, sy = 0.0f, sz = 0.0f, rho = 0.0f;
    visit([&](int, const float4 pj, const float4 vj) {
        const float dx = pi.x - pj.x;
        const float dy = pi.y - pj.y;
        const float dz = pi.z - pj.z;
        const float r2 = dist2<kFast>(dx, dy, dz);
        if (r2 <= k.h2) {
            const float t = poly_t<kFast>(r2, k);
            const float w = (t * t) * t;
            sx = acc<kFast>(sx, vj.x - vi.x, w);
            sy = acc<kFast>(sy, vj.y - vi.y, w);
            sz = acc<kFast>(sz, vj.z - vi.z, w);
            if (kCheck) rho += w;
        }
    });
    const float4 vc = make_float4(vi.x + k.cW * sx, vi.y + k.cW * sy, vi.z + k.cW * sz, 0.0f);
    const float4 xc = make_float4(pi.x, pi.y, pi.z, 0.0f);
    vout[i] = vc;
    xout[i] = xc;
    if (xn) *xn = xc;
    if (vn) *vn = vc;
    const float Ci = rho * k.c_rho - 1.0f;
    // a non-finite position (no pair passes the distance test) or a NaN
    // must fail the density test
    const bool finite = __builtin_isfinite(pi.x) && __builtin_isfinite(pi.y) && __builtin_isfinite(pi.z);
    return (finite && Ci == Ci) ? Ci : __builtin_inff();
}

// The next step's predict + cell key of slot i from the committed x and v,
// as k_pbf_key computes them from the stored arrays (the same floats, so the
// same keys); every lane of the wave calls it (histogram slots by wave runs).
__device__ __forceinline__ void next_key(int i, bool active, const float4 xn, const float4 vn, const PbfK& k,
                                         const NextKeys& nk) {
    int c = 0;
    if (active) {
        const Bounds b = bounds_of(k, k.scenes ? k.sid_out[i] : 0);
        c = pbf_cell(pbf_predict(xn, vn, k, b, nk.dt), k, b.yoff, 0);
        nk.key[i] = c;
    }
    const int sl = wave_atomic_slot(nk.counts, c, active);
    if (active) nk.slot[i] = sl;
}

template <bool kFast, bool kCheck>
__global__ __launch_bounds__(kBlock) void k_pbf_xsph(const float4* __restrict__ xa,
                                                      const float4* __restrict__ vt,
                                                      int* __restrict__ key2,
                                                      const int* __restrict__ cs, int n, PbfK k,
                                                      float4* __restrict__ vout,
                                                      float4* __restrict__ xout,
                                                      unsigned long long* __restrict__ masks,
                                                      const float4* __restrict__ xp,
                                                      int* __restrict__ maxd,
                                                      const int* __restrict__ gate, NextKeys nk) {
    if (gate && *gate == 0) return;  // uniform: before any barrier
    const int i = xcd_index();
    float Ci = -1e30f;
    float4 xn = make_float4(0.0f, 0.0f, 0.0f, 0.0f), vn = xn;
    if (i < n)
        Ci = xsph_body<kFast, kCheck>(i, xa, vt, k, vout, xout, [&](auto&& f) {
            for_each_neighbor<false, kWalkPairs>(i, n, key2, cs, masks, xp, xa, vt, k, f);
        }, &xn, &vn);
    if (nk.key) next_key(i, i < n, xn, vn, k, nk);
    if (kCheck) block_max_sum_store(Ci, Ci > 0.0f ? Ci : 0.0f, maxd);
}

template <bool kFast, bool kCheck>
__global__ __launch_bounds__(kTileT) void k_pbf_xsph_t(const float4* __restrict__ xa,
                                                        const float4* __restrict__ vt,
                                                        int* __restrict__ key2,
                                                        const int* __restrict__ cs, int n, PbfK k,
                                                        float4* __restrict__ vout,
                                                        float4* __restrict__ xout,
                                                        unsigned long long* __restrict__ masks,
                                                        const float4* __restrict__ xp,
                                                        int* __restrict__ maxd,
                                                        const int* __restrict__ gate,
                                                        const int4* __restrict__ tiles,
                                                        const int* __restrict__ ntiles, int lcap) {
    if (gate && *gate == 0) return;
    extern __shared__ float4 lds[];  // positions [lcap], then velocities [lcap]
    __shared__ TileShared S;
    float cmax = -1e30f, csum = 0.0f;
    const int nt = *ntiles;
    for (int t = tile_first(); t < nt; t += gridDim.x) {
        tile_stage(S, tiles[t], cs, k, xa, vt, lds, lds + lcap, lcap);
        const int i = S.i0 + (int)threadIdx.x;
        if (i < S.i1) {
            const float c = xsph_body<kFast, kCheck>(i, xa, vt, k, vout, xout, [&](auto&& f) {
                for_each_neighbor_tiled<false>(S, i, n, key2, masks, xp, xa, vt, lds, lds + lcap, k, f);
            });
            cmax = c > cmax ? c : cmax;
            csum += c > 0.0f ? c : 0.0f;
        }
        __syncthreads();
    }
    if (kCheck) block_max_sum_store(cmax, csum, maxd);
}

// ---------------------------------------------------------------------------
// Small scenes (fused arithmetic): the list walks of the plain step with kSub
// lanes per particle.  A scene of tens of thousands of particles fills a
// fraction of the GPU's wave slots, and each pass takes as long as one
// wave's walk of its longest list; kSub lanes walking every kSub-th group of
// eight entries shorten that walk kSub times.  The partial sums are combined
// in a fixed order ((s0 + s1) + (s2 + s3), the same in every lane), so the
// results are deterministic, but not in the oracle's order: only fused
// mode uses these kernels.  A particle whose list overflowed is walked by
// its first lane alone (the re-test path).
// ---------------------------------------------------------------------------
constexpr int kSub = kSubLanes;
static_assert(kSub == 4, "sub_sum combines four lanes");

// Walk the groups sub, sub + kSub, ... of the list of header hd
// (|N(i)| <= kNbrCap), calling f(j, cur[j], aux[j]) in list order within
// each group; the eight gathers of a group are in flight together.
template <class F>
__device__ __forceinline__ void visit_list_sub(const int4 hd, const int4* __restrict__ lst,
                                               const float4* __restrict__ cur, const float4* __restrict__ aux,
                                               int sub, F& f) {
    const int c = hd.w & 0xFFFF;
    const int n0 = (hd.w >> 16) & 0xFF, n01 = (int)((unsigned int)hd.w >> 24);
    auto slot = [&](int t, unsigned int e) { return (t < n0 ? hd.x : (t < n01 ? hd.y : hd.z)) + (int)e; };
    const float4 z = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    for (int g = sub; g * 8 < c; g += kSub) {
        const int4 q = lst[(size_t)g * 64];
        const int t = g * 8;
        const unsigned int w0 = (unsigned int)q.x, w1 = (unsigned int)q.y;
        const unsigned int w2 = (unsigned int)q.z, w3 = (unsigned int)q.w;
        const int j[8] = {slot(t, w0 & 0xFFFFu), slot(t + 1, w0 >> 16), slot(t + 2, w1 & 0xFFFFu),
                          slot(t + 3, w1 >> 16), slot(t + 4, w2 & 0xFFFFu), slot(t + 5, w2 >> 16),
                          slot(t + 6, w3 & 0xFFFFu), slot(t + 7, w3 >> 16)};
        float4 p[8], a[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            p[u] = at4(cur, j[u]);
            a[u] = aux ? at4(aux, j[u]) : z;
        }
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (t + u < c) f(j[u], p[u], a[u]);
    }
}

// the sum of v over the kSub lanes of a particle, (s0 + s1) + (s2 + s3) in every lane
__device__ __forceinline__ float sub_sum(float v) {
    v += __shfl_xor(v, 1, 64);
    v += __shfl_xor(v, 2, 64);
    return v;
}

// N(i) for lane sub of particle i: its share of the list, or the whole re-test
// walk for the first lane
template <class F>
__device__ __forceinline__ void for_each_neighbor_sub(int i, int sub, int n, int* __restrict__ key2,
                                                      const int* __restrict__ cs,
                                                      unsigned long long* __restrict__ nb,
                                                      const float4* __restrict__ xp, const float4* __restrict__ cur,
                                                      const float4* __restrict__ aux, const PbfK& k, F&& f) {
    const int4 hd = nbr_hdr(nb)[i];
    if ((hd.w & 0xFFFF) <= kNbrCap)
        visit_list_sub(hd, nbr_list(nb, n, i), cur, aux, sub, f);
    else if (sub == 0)
        for_each_neighbor<false>(i, n, key2, cs, nb, xp, cur, aux, k, f);
}

__global__ __launch_bounds__(kBlock) void k_pbf_lambda_sub(const float4* __restrict__ xa,
                                                            int* __restrict__ key2, const int* __restrict__ cs,
                                                            int n, PbfK k, float4* __restrict__ xl,
                                                            int* __restrict__ maxc,
                                                            unsigned long long* __restrict__ nb,
                                                            const float4* __restrict__ xp,
                                                            const int* __restrict__ gate) {
    if (gate && *gate == 0) return;  // uniform: before any barrier
    const int t = xcd_index();
    const int i = t / kSub, sub = t % kSub;
    LambdaAcc<true> acc(k);
    float4 pi = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (i < n) {
        pi = xa[i];
        for_each_neighbor_sub(i, sub, n, key2, cs, nb, xp, xa, nullptr, k,
                              [&](int, const float4 pj, const float4) { acc.add(pi, pj, k); });
    }
    // every lane takes part in the exchanges
    acc.rho = sub_sum(acc.rho);
    acc.a = sub_sum(acc.a);
    acc.gx = sub_sum(acc.gx);
    acc.gy = sub_sum(acc.gy);
    acc.gz = sub_sum(acc.gz);
    float Ci = -1e30f;
    if (i < n && sub == 0) Ci = acc.finish(pi, xl + i, k);
    block_max_store(Ci, maxc);
}

template <int kNS>
__global__ __launch_bounds__(kBlock) void k_pbf_deltap_sub(const float4* __restrict__ xl,
                                                            int* __restrict__ key2, const int* __restrict__ cs,
                                                            int n, PbfK k, float4* __restrict__ xa,
                                                            const float4* __restrict__ xold,
                                                            float4* __restrict__ vout, float dt,
                                                            unsigned long long* __restrict__ nb,
                                                            const float4* __restrict__ xp,
                                                            const int* __restrict__ gate) {
    if (gate && *gate == 0) return;
    const int t = xcd_index();
    const int i = t / kSub, sub = t % kSub;
    DeltapAcc<true, kNS> d(k);
    float4 pi = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (i < n) {
        pi = xl[i];
        const float li = pi.w;
        for_each_neighbor_sub(i, sub, n, key2, cs, nb, xp, xl, nullptr, k,
                              [&](int, const float4 pj, const float4) { d.add(pi, li, pj, k); });
    }
    const float ax = sub_sum(d.ax), ay = sub_sum(d.ay), az = sub_sum(d.az);
    if (i < n && sub == 0) deltap_commit(i, pi, ax, ay, az, d.step, k, xa, xold, vout, dt);
}

__global__ __launch_bounds__(kBlock) void k_pbf_xsph_sub(const float4* __restrict__ xa,
                                                          const float4* __restrict__ vt,
                                                          int* __restrict__ key2, const int* __restrict__ cs,
                                                          int n, PbfK k, float4* __restrict__ vout,
                                                          float4* __restrict__ xout,
                                                          unsigned long long* __restrict__ nb,
                                                          const float4* __restrict__ xp,
                                                          const int* __restrict__ gate, NextKeys nk) {
    if (gate && *gate == 0) return;
    const int t = xcd_index();
    const int i = t / kSub, sub = t % kSub;
    float4 pi = make_float4(0.0f, 0.0f, 0.0f, 0.0f), vi = pi;
    float sx = 0.0f, sy = 0.0f, sz = 0.0f;
    if (i < n) {
        pi = xa[i];
        vi = vt[i];
        for_each_neighbor_sub(i, sub, n, key2, cs, nb, xp, xa, vt, k, [&](int, const float4 pj, const float4 vj) {
            const float dx = pi.x - pj.x;
            const float dy = pi.y - pj.y;
            const float dz = pi.z - pj.z;
            const float r2 = dist2<true>(dx, dy, dz);
            if (r2 <= k.h2) {
                const float tt = poly_t<true>(r2, k);
                const float w = (tt * tt) * tt;
                sx = acc<true>(sx, vj.x - vi.x, w);
                sy = acc<true>(sy, vj.y - vi.y, w);
                sz = acc<true>(sz, vj.z - vi.z, w);
            }
        });
    }
    sx = sub_sum(sx);
    sy = sub_sum(sy);
    sz = sub_sum(sz);
    const float4 vn = make_float4(vi.x + k.cW * sx, vi.y + k.cW * sy, vi.z + k.cW * sz, 0.0f);
    const float4 xn = make_float4(pi.x, pi.y, pi.z, 0.0f);
    if (i < n && sub == 0) {
        vout[i] = vn;
        xout[i] = xn;
    }
    if (nk.key) {
        // the wave's 16 particles to lanes 0..15 (in slot order), so equal
        // keys of neighbouring slots share one histogram atomic
        const int lane = (int)(threadIdx.x & 63);
        const int src = (lane & 15) * kSub;
        const float4 xq = make_float4(__shfl(xn.x, src, 64), __shfl(xn.y, src, 64), __shfl(xn.z, src, 64), 0.0f);
        const float4 vq = make_float4(__shfl(vn.x, src, 64), __shfl(vn.y, src, 64), __shfl(vn.z, src, 64), 0.0f);
        const int iq = (t - lane) / kSub + (lane & 15);
        next_key(iq, lane < 16 && iq < n, xq, vq, k, nk);
    }
}

// The list build of a small fused scene with one lane per stencil layer
// (rows 3L..3L+2; the fourth lane of a particle idles).  A first walk over
// the layer's candidates counts the hits, notes the layer's first and last
// neighbour and sums lambda; the lanes exchange their counts, and a second
// walk stores each hit's 16-bit entry at its place in the list (the same
// entries in the same order as the one-lane build) with two-byte stores.
// The last layer's lane zeroes the places past the count in the last group
// (a walk decodes them, to slot b2).
__global__ __launch_bounds__(kBlock) void k_pbf_lambda_build_sub(const float4* __restrict__ xp,
                                                                  const int* __restrict__ key2,
                                                                  const int* __restrict__ cs, int n, PbfK k,
                                                                  float4* __restrict__ xl,
                                                                  int* __restrict__ maxc,
                                                                  unsigned long long* __restrict__ nb,
                                                                  const int* __restrict__ gate) {
    if (gate && *gate == 0) return;  // uniform: before any barrier
    const int t = xcd_index();
    const int i = t / kSub, layer = t % kSub;
    const bool mine = i < n && layer < 3;
    LambdaAcc<true> acc(k);
    float4 pi = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    int cnt = 0, first = -1, last = -1;
    int cy = 0, cx = 0, cz = 0;
    if (mine) {
        key_cell(key2[i] & ~kOvf, k, cy, cx, cz);
        pi = xp[i];
    }
    const int z0 = cz > 0 ? cz - 1 : 0;
    const int z1 = cz + 1 < k.gZ ? cz + 1 : k.gZ - 1;
    const int y = cy + layer - 1;
    const bool yok = mine && y >= 0 && y < k.gY;
    // the layer's candidate rows (x - 1, x, x + 1), ascending slots; a row of
    // at most 64 slots keeps its hits as a mask (row_m, from row_j0) for the
    // second walk, a longer one is tested again there
    unsigned long long row_m[3] = {0ull, 0ull, 0ull};
    int row_j0[3] = {0, 0, 0}, row_j1[3] = {0, 0, 0};
    if (yok) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const int x = cx + d - 1;
            if (x < 0 || x >= k.gX) continue;
            const int row = row_key(y, x, 0, k);
            row_j0[d] = cs[row + z0];
            row_j1[d] = cs[row + z1 + 1];
        }
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const int j0 = row_j0[d], j1 = row_j1[d];
        unsigned long long m = 0ull;
        int j = j0;
        for (; j + 4 <= j1; j += 4) {  // four candidates' loads in flight
            const float3 c[4] = {xyz_at(xp, j), xyz_at(xp, j + 1), xyz_at(xp, j + 2), xyz_at(xp, j + 3)};
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (within_h<true>(pi, c[u], k)) {
                    const int jj = j + u;
                    first = first < 0 ? jj : first;
                    last = jj;
                    ++cnt;
                    acc.add(pi, make_float4(c[u].x, c[u].y, c[u].z, 0.0f), k);
                    if (jj - j0 < 64) m |= 1ull << (jj - j0);
                }
        }
        for (; j < j1; ++j) {
            const float3 c = xyz_at(xp, j);
            if (within_h<true>(pi, c, k)) {
                first = first < 0 ? j : first;
                last = j;
                ++cnt;
                acc.add(pi, make_float4(c.x, c.y, c.z, 0.0f), k);
                if (j - j0 < 64) m |= 1ull << (j - j0);
            }
        }
        row_m[d] = m;
    }
    // the particle's three layers (lanes base .. base + 2)
    const int base = (int)(threadIdx.x & 63) & ~(kSub - 1);
    const int c0 = __shfl(cnt, base, 64), c1 = __shfl(cnt, base + 1, 64), c2 = __shfl(cnt, base + 2, 64);
    const int f0 = __shfl(first, base, 64), f1 = __shfl(first, base + 1, 64), f2 = __shfl(first, base + 2, 64);
    const int wide = (mine && first >= 0 && last - first > 0xFFFF) ? 1 : 0;
    const int anywide = __shfl(wide, base, 64) | __shfl(wide, base + 1, 64) | __shfl(wide, base + 2, 64);
    const int c = c0 + c1 + c2;
    const bool retest = c > kNbrCap || anywide;
    acc.rho = sub_sum(acc.rho);
    acc.a = sub_sum(acc.a);
    acc.gx = sub_sum(acc.gx);
    acc.gy = sub_sum(acc.gy);
    acc.gz = sub_sum(acc.gz);
    if (mine && !retest) {
        int4* __restrict__ lst = nbr_list(nb, n, i);
        // entry p of the list: group p / 8, 16-bit place p % 8 of the int4
        auto place = [&](int p) {
            return reinterpret_cast<unsigned short*>(lst + (size_t)(p >> 3) * 64) + (p & 7);
        };
        int p = layer == 0 ? 0 : (layer == 1 ? c0 : c0 + c1);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const int j0 = row_j0[d], j1 = row_j1[d];
            if (j1 - j0 <= 64) {
                for (unsigned long long m = row_m[d]; m; m &= m - 1ull)
                    *place(p++) = (unsigned short)(j0 + low_bit(m) - first);
            } else {
                for (int j = j0; j < j1; ++j)
                    if (within_h<true>(pi, xyz_at(xp, j), k)) *place(p++) = (unsigned short)(j - first);
            }
        }
        if (layer == 2)
            for (int q = c; (q & 7) != 0; ++q) *place(q) = 0;
    }
    float Ci = -1e30f;
    if (i < n && layer == 0) {
        Ci = acc.finish(pi, xl + i, k);
        const int b0 = f0 < 0 ? i : f0, b1 = f1 < 0 ? i : f1, b2 = f2 < 0 ? i : f2;
        const int w = retest ? kNbrRetest : (c | (c0 << 16) | ((c0 + c1) << 24));
        nbr_hdr(nb)[i] = make_int4(b0, b1, b2, w);
    }
    block_max_store(Ci, maxc);
}

hipError_t launch_lambda_build_sub(const float4* xp, const int* key2, const int* cs, int n, PbfK k, float4* xl,
                                   int* maxc, unsigned long long* masks, const int* gate, hipStream_t s) {
    launch_kernel(k_pbf_lambda_build_sub, dim3(grid_for(n * kSub)), dim3(kBlock), 0, s, xp, key2, cs, n, k, xl,
                  maxc, masks, gate);
    return hipGetLastError();
}

hipError_t launch_lambda_sub(const float4* xa, int* key2, const int* cs, int n, PbfK k, float4* xl, int* maxc,
                             unsigned long long* masks, const float4* xp, const int* gate, hipStream_t s) {
    launch_kernel(k_pbf_lambda_sub, dim3(grid_for(n * kSub)), dim3(kBlock), 0, s, xa, key2, cs, n, k, xl, maxc,
                  masks, xp, gate);
    return hipGetLastError();
}

hipError_t launch_deltap_sub(const float4* xl, int* key2, const int* cs, int n, PbfK k, float4* xa,
                             const float4* xold, float4* vout, float dt, unsigned long long* masks,
                             const float4* xp, const int* gate, hipStream_t s) {
    const dim3 g(grid_for(n * kSub)), b(kBlock);
    if (k.n_scorr == 4)
        launch_kernel((k_pbf_deltap_sub<4>), g, b, 0, s, xl, key2, cs, n, k, xa, xold, vout, dt, masks, xp, gate);
    else
        launch_kernel((k_pbf_deltap_sub<0>), g, b, 0, s, xl, key2, cs, n, k, xa, xold, vout, dt, masks, xp, gate);
    return hipGetLastError();
}

hipError_t launch_xsph_sub(const float4* xa, const float4* vt, int* key2, const int* cs, int n, PbfK k,
                           float4* vout, float4* xout, unsigned long long* masks, const float4* xp,
                           const int* gate, hipStream_t s, const NextKeys& nk) {
    launch_kernel(k_pbf_xsph_sub, dim3(grid_for(n * kSub)), dim3(kBlock), 0, s, xa, vt, key2, cs, n, k, vout, xout,
                  masks, xp, gate, nk);
    return hipGetLastError();
}

hipError_t launch_pbf_xsph(const float4* xa, const float4* vt, int* key2, const int* cs, int n,
                           PbfK k, float4* vout, float4* xout, unsigned long long* masks,
                           const float4* xp, int* maxd, const int* gate, const TileArgs& ta, hipStream_t s,
                           const NextKeys* next) {
    const NextKeys nk = next ? *next : NextKeys();
    if (ta.tiles) {
        if (nk.key) return hipErrorInvalidValue;
        const dim3 g(ta.grid), b(kTileT);
        const size_t sh = 2 * (size_t)ta.lcap * sizeof(float4);
        if (k.exact && maxd)
            launch_kernel((k_pbf_xsph_t<false, true>), g, b, sh, s, xa, vt, key2, cs, n, k, vout, xout, masks,
                               xp, maxd, gate, ta.tiles, ta.ntiles, ta.lcap);
        else if (k.exact)
            launch_kernel((k_pbf_xsph_t<false, false>), g, b, sh, s, xa, vt, key2, cs, n, k, vout, xout, masks,
                               xp, maxd, gate, ta.tiles, ta.ntiles, ta.lcap);
        else if (maxd)
            launch_kernel((k_pbf_xsph_t<true, true>), g, b, sh, s, xa, vt, key2, cs, n, k, vout, xout, masks,
                               xp, maxd, gate, ta.tiles, ta.ntiles, ta.lcap);
        else
            launch_kernel((k_pbf_xsph_t<true, false>), g, b, sh, s, xa, vt, key2, cs, n, k, vout, xout, masks,
                               xp, maxd, gate, ta.tiles, ta.ntiles, ta.lcap);
        return hipGetLastError();
    }
    if (ta.sub > 1 && !k.exact && !maxd)
        return launch_xsph_sub(xa, vt, key2, cs, n, k, vout, xout, masks, xp, gate, s, nk);
    const dim3 g(grid_for(n, ta.block)), b(ta.block);
    if (k.exact && maxd)
        launch_kernel((k_pbf_xsph<false, true>), g, b, 0, s, xa, vt, key2, cs, n, k, vout, xout, masks, xp, maxd, gate,
                      nk);
    else if (k.exact)
        launch_kernel((k_pbf_xsph<false, false>), g, b, 0, s, xa, vt, key2, cs, n, k, vout, xout, masks, xp, maxd, gate,
                      nk);
    else if (maxd)
        launch_kernel((k_pbf_xsph<true, true>), g, b, 0, s, xa, vt, key2, cs, n, k, vout, xout, masks, xp, maxd, gate,
                      nk);
    else
        launch_kernel((k_pbf_xsph<true, false>), g, b, 0, s, xa, vt, key2, cs, n, k, vout, xout, masks, xp, maxd, gate,
                      nk);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Mixed sand + fluid over static solids.  One cell-ordered array holds every
// phase (k.ph_out[slot]: 0 fluid, 1 sand, 2 solid; a neighbour's phase is read
// from w of its old, predicted or corrected position) and shares the neighbour
// sets.  Fluid: density over every neighbour, the other-particle gradient term
// over fluid neighbours only, lambda = 0 for sand and solids, delta-p as the
// PBF pass.  Sand: the contact + friction projection of simulate_sand
// (src/Simulate.cpp:240-300) against sand (half weight, friction), fluid
// (half weight, contact only) and solid (full weight, friction) neighbours
// within d, friction against the step's old positions.  Solids do not move.
// The sand arithmetic is IEEE in both arithmetic modes.
// ---------------------------------------------------------------------------
// xoj_of(): j's old position (its w: j's phase), asked for only for a contact
template <class XoJ>
__device__ __forceinline__ void mix_sand_pair(const float4 pi, const float4 xoi, const float4 pj, XoJ&& xoj_of,
                                              const PbfK& k, float& dpx, float& dpy, float& dpz) {
    float ijx = pi.x - pj.x, ijy = pi.y - pj.y, ijz = pi.z - pj.z;
    // most neighbours within h are beyond d: those whose squared length is
    // well above d^2 skip the square root (its result would exceed d)
    const float ij2 = dot3(ijx, ijy, ijz, ijx, ijy, ijz);
    if (ij2 > k.mix_d2_far) return;
    // |ij| == 0 -> ij = (0, 1e-5, 0) (src/Simulate.cpp:236-239); the length
    // is computed again only then
    float len = sqrtf(ij2);
    if (len == 0.0f) {
        ijx = 0.0f;
        ijy = 0.00001f;
        ijz = 0.0f;
        len = sqrtf(dot3(ijx, ijy, ijz, ijx, ijy, ijz));
    }
    if (len > k.mix_d) return;
    // j's old position and phase (w)
    const float4 xoj = xoj_of();
    const int jph = (int)xoj.w;
    float sc;
    if (jph == 2) {
        sc = k.mix_collision * (len - k.mix_d);
    } else {
        sc = k.mix_collision * 0.5f;
        sc = sc * (len - k.mix_d);
    }
    const float tx = (sc * ijx) / len, ty = (sc * ijy) / len, tz = (sc * ijz) / len;
    dpx = dpx - tx;
    dpy = dpy - ty;
    dpz = dpz - tz;
    if (jph == 0) return;  // fluid neighbour: contact only
    const float dd = sqrtf(dot3(tx, ty, tz, tx, ty, tz));
    const float xidx = (pi.x - tx) - xoi.x, xidy = (pi.y - ty) - xoi.y, xidz = (pi.z - tz) - xoi.z;
    float xjdx, xjdy, xjdz, nx, ny, nz;
    if (jph == 1) {
        xjdx = (pj.x + tx) - xoj.x;
        xjdy = (pj.y + ty) - xoj.y;
        xjdz = (pj.z + tz) - xoj.z;
        nx = (pi.x - tx) - (pj.x + tx);
        ny = (pi.y - ty) - (pj.y + ty);
        nz = (pi.z - tz) - (pj.z + tz);
    } else {
        xjdx = 0.0f;
        xjdy = 0.0f;
        xjdz = 0.0f;
        nx = (pi.x - tx) - pj.x;
        ny = (pi.y - ty) - pj.y;
        nz = (pi.z - tz) - pj.z;
    }
    const float inv = 1.0f / sqrtf(dot3(nx, ny, nz, nx, ny, nz));
    nx = nx * inv;
    ny = ny * inv;
    nz = nz * inv;
    const float rx = xidx - xjdx, ry = xidy - xjdy, rz = xidz - xjdz;
    const float rd = dot3(rx, ry, rz, nx, ny, nz);
    const float xtx = rx - rd * nx, xty = ry - rd * ny, xtz = rz - rd * nz;
    const float lt = sqrtf(dot3(xtx, xty, xtz, xtx, xty, xtz));
    if ((dd * k.mix_mu_s) > (lt + 1e-9f)) {
        dpx = dpx - k.mix_friction * xtx;
        dpy = dpy - k.mix_friction * xty;
        dpz = dpz - k.mix_friction * xtz;
    } else {
        const float q = (k.mix_mu_k * dd) / (lt + 1e-9f);
        const float ratio = q < 1.0f ? q : 1.0f;
        dpx = dpx - (k.mix_friction * xtx) * ratio;
        dpy = dpy - (k.mix_friction * xty) * ratio;
        dpz = dpz - (k.mix_friction * xtz) * ratio;
    }
}

template <bool kBuild, bool kFast>
__global__ __launch_bounds__(kBlock) void k_mix_lambda(const float4* __restrict__ xa, int* __restrict__ key2,
                                                       const int* __restrict__ cs, int n, PbfK k,
                                                       float4* __restrict__ xl,
                                                       unsigned long long* __restrict__ masks,
                                                       const float4* __restrict__ xp) {
    const int i = xcd_index();
    if (i >= n) return;
    const float4 pi = xa[i];
    const int ph = k.ph_out[i];
    if (ph != 0) {
        // no constraint; a building pass still writes the lists of sand (the
        // correction pass walks them), and an empty one for a solid, whose
        // list no pass reads
        if (kBuild && ph == 1)
            for_each_neighbor<true>(i, n, key2, cs, masks, xp, xa, nullptr, k, [](int, const float4, const float4) {});
        else if (kBuild)
            nbr_hdr(masks)[i] = make_int4(i, i, i, 0);
        xl[i] = make_float4(pi.x, pi.y, pi.z, 0.0f);
        return;
    }
    MixLambdaAcc<kFast> m;
    for_each_neighbor<kBuild, kWalkWide>(i, n, key2, cs, masks, xp, xa, nullptr, k,
                                         [&](int, const float4 pj, const float4) { m.add(pi, pj, k); });
    xl[i] = make_float4(pi.x, pi.y, pi.z, m.lambda(k));
}

hipError_t launch_mix_lambda(const float4* xa, int* key2, const int* cs, int n, PbfK k, float4* xl,
                             unsigned long long* masks, const float4* xp, int build, int stage, hipStream_t s) {
    const dim3 g(grid_for(n)), b(kBlock);
    if (build && stage && xa == xp && n > 0 && k.ph_out) {
        if (k.exact)
            launch_kernel((k_mix_lambda_build_ws<false>), g, b, 0, s, xp, key2, cs, n, k, xl, (int*)nullptr,
                          masks, (const int*)nullptr);
        else
            launch_kernel((k_mix_lambda_build_ws<true>), g, b, 0, s, xp, key2, cs, n, k, xl, (int*)nullptr,
                          masks, (const int*)nullptr);
        return hipGetLastError();
    }
    if (build && k.exact)
        launch_kernel((k_mix_lambda<true, false>), g, b, 0, s, xa, key2, cs, n, k, xl, masks, xp);
    else if (build)
        launch_kernel((k_mix_lambda<true, true>), g, b, 0, s, xa, key2, cs, n, k, xl, masks, xp);
    else if (k.exact)
        launch_kernel((k_mix_lambda<false, false>), g, b, 0, s, xa, key2, cs, n, k, xl, masks, xp);
    else
        launch_kernel((k_mix_lambda<false, true>), g, b, 0, s, xa, key2, cs, n, k, xl, masks, xp);
    return hipGetLastError();
}

constexpr int kMixContacts = 16;  // contacts of a sand particle held for the projection loop

template <bool kFast, int kNS>
__global__ __launch_bounds__(kBlock) void k_mix_correct(const float4* __restrict__ xl, int* __restrict__ key2,
                                                        const int* __restrict__ cs, int n, PbfK k,
                                                        float4* __restrict__ xa, const float4* __restrict__ xold,
                                                        float4* __restrict__ vout, float dt,
                                                        unsigned long long* __restrict__ masks,
                                                        const float4* __restrict__ xp) {
    // contacts of a sand lane, in list order: contact[u * kBlock + thread]
    __shared__ int contact[kMixContacts * kBlock];
    const int i = xcd_index();
    if (i >= n) return;
    const int ph = k.ph_out[i];
    const float4 pi = xl[i];
    float4 r = make_float4(pi.x, pi.y, pi.z, (float)ph);
    if (ph == 1) {
        const float4 xoi = xold[i];
        float dpx = 0.0f, dpy = 0.0f, dpz = 0.0f;
        // A contact (within d) is rare per neighbour but, over a wave, some
        // lane has one at almost every step of the walk: the walk only
        // collects the neighbours that can be within d (mix_d2_far), and a
        // second loop, as long as the wave's largest contact count, projects
        // them in list order (the same sums in the same order).
        int nc = 0;
        for_each_neighbor<false, kWalkWide>(i, n, key2, cs, masks, xp, xl, nullptr, k, [&](int j, const float4 pj, const float4) {
            if (j == i) return;
            const float ijx = pi.x - pj.x, ijy = pi.y - pj.y, ijz = pi.z - pj.z;
            if (dot3(ijx, ijy, ijz, ijx, ijy, ijz) > k.mix_d2_far) return;
            if (nc < kMixContacts) contact[nc * kBlock + (int)threadIdx.x] = j;
            ++nc;
        });
        if (nc <= kMixContacts) {
            // four contacts at a time: their slots, positions and old
            // positions requested together, then projected in list order
            // (places past nc read slot i and are not projected)
            for (int u = 0; u < nc; u += 4) {
                int jj[4];
                float4 pj[4], xoj[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) jj[q] = u + q < nc ? contact[(u + q) * kBlock + (int)threadIdx.x] : i;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    pj[q] = xl[jj[q]];
                    xoj[q] = xold[jj[q]];
                }
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    if (u + q < nc) mix_sand_pair(pi, xoi, pj[q], [&] { return xoj[q]; }, k, dpx, dpy, dpz);
            }
        } else {
            // more than the buffer holds: the walk again, every neighbour
            for_each_neighbor<false>(i, n, key2, cs, masks, xp, xl, nullptr, k,
                                     [&](int j, const float4 pj, const float4) {
                if (j == i) return;
                mix_sand_pair(pi, xoi, pj, [&] { return xold[j]; }, k, dpx, dpy, dpz);
            });
        }
        r.x = rclamp(pi.x + dpx, k.lo[0], k.hi[0]);
        r.y = rclamp(pi.y + dpy, k.lo[1], k.hi[1]);
        r.z = rclamp(pi.z + dpz, k.lo[2], k.hi[2]);
    } else if (ph == 0) {
        // the plain delta-p sums (lambda_j = 0 for sand and solids)
        const float li = pi.w;
        DeltapAcc<kFast, kNS> d(k);
        for_each_neighbor<false, kWalkWide>(i, n, key2, cs, masks, xp, xl, nullptr, k,
                                            [&](int, const float4 pj, const float4) { d.add(pi, li, pj, k); });
        r.x = rclamp(pi.x + d.ax * d.step, k.lo[0], k.hi[0]);
        r.y = rclamp(pi.y + d.ay * d.step, k.lo[1], k.hi[1]);
        r.z = rclamp(pi.z + d.az * d.step, k.lo[2], k.hi[2]);
    }
    xa[i] = r;
    if (vout) {
        const float4 xo = xold[i];
        vout[i] = make_float4((r.x - xo.x) / dt, (r.y - xo.y) / dt, (r.z - xo.z) / dt, 0.0f);
    }
}

hipError_t launch_mix_correct(const float4* xl, int* key2, const int* cs, int n, PbfK k, float4* xa,
                              const float4* xold, float4* vout, float dt, unsigned long long* masks,
                              const float4* xp, hipStream_t s) {
    const dim3 g(grid_for(n)), b(kBlock);
    if (k.exact)
        launch_kernel((k_mix_correct<false, 0>), g, b, 0, s, xl, key2, cs, n, k, xa, xold, vout, dt, masks, xp);
    else if (k.n_scorr == 4)
        launch_kernel((k_mix_correct<true, 4>), g, b, 0, s, xl, key2, cs, n, k, xa, xold, vout, dt, masks, xp);
    else
        launch_kernel((k_mix_correct<true, 0>), g, b, 0, s, xl, key2, cs, n, k, xa, xold, vout, dt, masks, xp);
    return hipGetLastError();
}

template <bool kFast, bool kCheck>
__global__ __launch_bounds__(kBlock) void k_mix_xsph(const float4* __restrict__ xa, const float4* __restrict__ vt,
                                                     int* __restrict__ key2, const int* __restrict__ cs, int n,
                                                     PbfK k, float4* __restrict__ vout, float4* __restrict__ xout,
                                                     unsigned long long* __restrict__ masks,
                                                     const float4* __restrict__ xp, int* __restrict__ maxd) {
    const int i = xcd_index();
    float Ci = -1e30f;
    if (i < n) {
        const float4 pi = xa[i];
        const float4 vi = vt[i];
        if (k.ph_out[i] != 0) {
            vout[i] = vi;
            xout[i] = make_float4(pi.x, pi.y, pi.z, 0.0f);
        } else {
            float sx = 0.0f, sy = 0.0f, sz = 0.0f, rho = 0.0f;
            for_each_neighbor<false, kWalkPairs>(i, n, key2, cs, masks, xp, xa, vt, k,
                                     [&](int j, const float4 pj, const float4 vj) {
                const float dx = pi.x - pj.x;
                const float dy = pi.y - pj.y;
                const float dz = pi.z - pj.z;
                const float r2 = dist2<kFast>(dx, dy, dz);
                if (r2 <= k.h2) {
                    const float t = poly_t<kFast>(r2, k);
                    const float w = (t * t) * t;
                    rho += w;
                    if (pj.w == 0.0f) {
                        sx = acc<kFast>(sx, vj.x - vi.x, w);
                        sy = acc<kFast>(sy, vj.y - vi.y, w);
                        sz = acc<kFast>(sz, vj.z - vi.z, w);
                    }
                }
            });
            vout[i] = make_float4(vi.x + k.cW * sx, vi.y + k.cW * sy, vi.z + k.cW * sz, 0.0f);
            xout[i] = make_float4(pi.x, pi.y, pi.z, 0.0f);
            const float c = rho * k.c_rho - 1.0f;
            const bool finite = __builtin_isfinite(pi.x) && __builtin_isfinite(pi.y) && __builtin_isfinite(pi.z);
            Ci = (finite && c == c) ? c : __builtin_inff();
        }
    }
    if (kCheck) block_max_sum_store(Ci, Ci > 0.0f ? Ci : 0.0f, maxd);
}

hipError_t launch_mix_xsph(const float4* xa, const float4* vt, int* key2, const int* cs, int n, PbfK k,
                           float4* vout, float4* xout, unsigned long long* masks, const float4* xp, int* maxd,
                           hipStream_t s) {
    const dim3 g(grid_for(n)), b(kBlock);
    if (k.exact && maxd)
        launch_kernel((k_mix_xsph<false, true>), g, b, 0, s, xa, vt, key2, cs, n, k, vout, xout, masks, xp, maxd);
    else if (k.exact)
        launch_kernel((k_mix_xsph<false, false>), g, b, 0, s, xa, vt, key2, cs, n, k, vout, xout, masks, xp, maxd);
    else if (maxd)
        launch_kernel((k_mix_xsph<true, true>), g, b, 0, s, xa, vt, key2, cs, n, k, vout, xout, masks, xp, maxd);
    else
        launch_kernel((k_mix_xsph<true, false>), g, b, 0, s, xa, vt, key2, cs, n, k, vout, xout, masks, xp, maxd);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// learned warm start: features, training targets, fallback decision
// ---------------------------------------------------------------------------
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));

// Features of particle i at the predicted positions x*, over N(i):
//   f[0]        C_i = rho_i / rho0 - 1 (poly6, self included)
//   f[1..3]     (x*_i - x_i) / h, the predicted displacement of the step
//   f[4+3b..]   b = 0..7: mean offset (x*_j - x*_i) / h over j != i with
//               b = min(7, (int)(8 |d| / h)) (0 when the bin is empty)
//   f[28]       |N(i) \ i| / 32
//   f[29]       mean |d| / h over N(i) \ i
//   f[30]       mean |d|^2 / h^2 over N(i) \ i
//   f[31]       min(1, distance to the nearest wall of the clamp box / h)
// kLdsBins: this thread's 8 bin accumulators (x, y, z, count) live in LDS,
// bin b at bins[b * kBlock]: one read-modify-write per neighbour instead of a
// one-hot update of 32 registers (the same additions in the same order)
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) f32x4 lds_float4;

template <bool kFast, bool kRecord, bool kLdsBins, class Visit>
__device__ __forceinline__ void features_body(int i, const float4* __restrict__ xp, const float4* __restrict__ xold,
                                              const PbfK& k, const float* __restrict__ mu,
                                              const float* __restrict__ inv_sigma, bf16x8_t* __restrict__ fbf,
                                              float* __restrict__ frec, int rec_stride, int rec_off, float inv_h,
                                              Visit&& visit, lds_float4* bins = nullptr) {
    const float4 pi = xp[i];
    float rho = 0.0f, cnt = 0.0f, sr = 0.0f, sr2 = 0.0f;
    float bx[8], by[8], bz[8], bn[8];
#pragma unroll
    for (int b = 0; b < 8; ++b) bx[b] = by[b] = bz[b] = bn[b] = 0.0f;
    if constexpr (kLdsBins) {
#pragma unroll
        for (int b = 0; b < 8; ++b) bins[b * kBlock] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    }
    visit([&](int j, const float4 pj, const float4) {
        const float dx = pj.x - pi.x;
        const float dy = pj.y - pi.y;
        const float dz = pj.z - pi.z;
        const float r2 = dx * dx + dy * dy + dz * dz;
        const float t = poly_t<kFast>(r2, k);
        rho = acc<kFast>(rho, t * t, t);
        if (j != i) {
            const float q = sqrtf(r2) * inv_h;
            int b = (int)(q * 8.0f);
            b = b < 7 ? b : 7;
            if constexpr (kLdsBins) {
                lds_float4* __restrict__ p = bins + b * kBlock;
                f32x4 v = *p;
                v.x += dx;
                v.y += dy;
                v.z += dz;
                v.w += 1.0f;
                *p = v;
            } else {
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const float on = (b == u) ? 1.0f : 0.0f;
                    bx[u] += on * dx;
                    by[u] += on * dy;
                    bz[u] += on * dz;
                    bn[u] += on;
                }
            }
            cnt += 1.0f;
            sr += q;
            sr2 += q * q;
        }
    });
    if constexpr (kLdsBins) {
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const f32x4 v = bins[b * kBlock];
            bx[b] = v.x;
            by[b] = v.y;
            bz[b] = v.z;
            bn[b] = v.w;
        }
    }
    float f[kFeat];
    f[0] = rho * k.c_rho - 1.0f;
    const float4 xo = xold[i];
    f[1] = (pi.x - xo.x) * inv_h;
    f[2] = (pi.y - xo.y) * inv_h;
    f[3] = (pi.z - xo.z) * inv_h;
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        const float s = bn[b] > 0.0f ? inv_h / bn[b] : 0.0f;
        f[4 + 3 * b] = bx[b] * s;
        f[5 + 3 * b] = by[b] * s;
        f[6 + 3 * b] = bz[b] * s;
    }
    f[28] = cnt * (1.0f / 32.0f);
    f[29] = cnt > 0.0f ? sr / cnt : 0.0f;
    f[30] = cnt > 0.0f ? sr2 / cnt : 0.0f;
    const Bounds bb = bounds_of(k, k.scenes ? k.sid_out[i] : 0);
    float wall = pi.x - bb.lo[0];
    wall = fminf(wall, bb.hi[0] - pi.x);
    wall = fminf(wall, pi.y - bb.lo[1]);
    wall = fminf(wall, bb.hi[1] - pi.y);
    wall = fminf(wall, pi.z - bb.lo[2]);
    wall = fminf(wall, bb.hi[2] - pi.z);
    f[31] = fminf(1.0f, wall * inv_h);
    if (kRecord) {
        const int p0 = (rec_stride - rec_off % rec_stride) % rec_stride;
        if (i < p0 || (i - p0) % rec_stride != 0) return;
        float4* row = reinterpret_cast<float4*>(frec + (size_t)((i - p0) / rec_stride) * kFeat);
#pragma unroll
        for (int q = 0; q < kFeat / 4; ++q) row[q] = make_float4(f[4 * q], f[4 * q + 1], f[4 * q + 2], f[4 * q + 3]);
    } else {
#pragma unroll
        for (int q = 0; q < kFeat / 8; ++q) {
            bf16x8_t v;
#pragma unroll
            for (int e = 0; e < 8; ++e) v[e] = (__bf16)((f[8 * q + e] - mu[8 * q + e]) * inv_sigma[8 * q + e]);
            fbf[(size_t)i * (kFeat / 8) + q] = v;
        }
    }
}

template <bool kFast, bool kRecord>
__global__ __launch_bounds__(kBlock) void k_pbf_features(
    int* __restrict__ key2, const int* __restrict__ cs, int n, PbfK k,
    unsigned long long* __restrict__ masks, const float4* __restrict__ xp,
    const float4* __restrict__ xold, const float* __restrict__ mu, const float* __restrict__ inv_sigma,
    bf16x8_t* __restrict__ fbf, float* __restrict__ frec, int rec_stride, int rec_off, float inv_h) {
    __shared__ float4 bins[8 * kBlock];  // blockDim.x <= kBlock
    const int i = xcd_index();
    if (i >= n) return;
    features_body<kFast, kRecord, true>(i, xp, xold, k, mu, inv_sigma, fbf, frec, rec_stride, rec_off, inv_h,
                                  [&](auto&& f) {
        for_each_neighbor<true>(i, n, key2, cs, masks, xp, xp, nullptr, k, f);
    }, (lds_float4*)bins + threadIdx.x);
}

template <bool kFast>
__global__ __launch_bounds__(kTileT) void k_pbf_features_t(
    int* __restrict__ key2, const int* __restrict__ cs, int n, PbfK k,
    unsigned long long* __restrict__ masks, const float4* __restrict__ xp,
    const float4* __restrict__ xold, const float* __restrict__ mu, const float* __restrict__ inv_sigma,
    bf16x8_t* __restrict__ fbf, float inv_h, const int4* __restrict__ tiles, const int* __restrict__ ntiles,
    int lcap) {
    extern __shared__ float4 lds[];
    __shared__ TileShared S;
    const int nt = *ntiles;
    for (int t = tile_first(); t < nt; t += gridDim.x) {
        tile_stage(S, tiles[t], cs, k, xp, nullptr, lds, nullptr, lcap);
        const int i = S.i0 + (int)threadIdx.x;
        if (i < S.i1)
            features_body<kFast, false, false>(i, xp, xold, k, mu, inv_sigma, fbf, nullptr, 1, 0, inv_h, [&](auto&& f) {
                for_each_neighbor_tiled<true>(S, i, n, key2, masks, xp, xp, nullptr, lds, nullptr, k, f);
            });
        __syncthreads();
    }
}

hipError_t launch_pbf_features(int* key2, const int* cs, int n, PbfK k, unsigned long long* masks,
                               const float4* xp, const float4* xold, const float* mu,
                               const float* inv_sigma, void* fbf, float* frec, int rec_stride,
                               int rec_off, const TileArgs& ta, hipStream_t s) {
    const dim3 g(grid_for(n, ta.block)), b(ta.block);
    const float inv_h = 1.0f / sqrtf(k.h2);
    if (!frec && ta.tiles) {
        const size_t sh = (size_t)ta.lcap * sizeof(float4);
        if (k.exact)
            launch_kernel((k_pbf_features_t<false>), dim3(ta.grid), dim3(kTileT), sh, s, key2, cs, n, k, masks,
                               xp, xold, mu, inv_sigma, (bf16x8_t*)fbf, inv_h, ta.tiles, ta.ntiles, ta.lcap);
        else
            launch_kernel((k_pbf_features_t<true>), dim3(ta.grid), dim3(kTileT), sh, s, key2, cs, n, k, masks,
                               xp, xold, mu, inv_sigma, (bf16x8_t*)fbf, inv_h, ta.tiles, ta.ntiles, ta.lcap);
        return hipGetLastError();
    }
    if (frec) {
        if (rec_stride < 1) return hipErrorInvalidValue;
        launch_kernel((k_pbf_features<false, true>), g, b, 0, s, key2, cs, n, k, masks, xp, xold, mu,
                           inv_sigma, (bf16x8_t*)fbf, frec, rec_stride, rec_off, inv_h);
    } else if (k.exact) {
        launch_kernel((k_pbf_features<false, false>), g, b, 0, s, key2, cs, n, k, masks, xp, xold, mu,
                           inv_sigma, (bf16x8_t*)fbf, frec, 1, 0, inv_h);
    } else {
        launch_kernel((k_pbf_features<true, false>), g, b, 0, s, key2, cs, n, k, masks, xp, xold, mu,
                           inv_sigma, (bf16x8_t*)fbf, frec, 1, 0, inv_h);
    }
    return hipGetLastError();
}

__global__ __launch_bounds__(kBlock) void k_pbf_record_target(const float4* __restrict__ xa,
                                                               const float4* __restrict__ xp, int n,
                                                               int rec_stride, int rec_off,
                                                               float* __restrict__ tgt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int p0 = (rec_stride - rec_off % rec_stride) % rec_stride;
    const int p = p0 + i * rec_stride;
    if (p >= n) return;
    const float4 a = xa[p], b = xp[p];
    tgt[(size_t)i * 3 + 0] = a.x - b.x;
    tgt[(size_t)i * 3 + 1] = a.y - b.y;
    tgt[(size_t)i * 3 + 2] = a.z - b.z;
}

hipError_t launch_pbf_record_target(const float4* xa, const float4* xp, int n, int rec_stride,
                                    int rec_off, float* tgt, hipStream_t s) {
    if (rec_stride < 1) return hipErrorInvalidValue;
    const int rows = n / rec_stride + 1;
    launch_kernel(k_pbf_record_target, dim3((rows + kBlock - 1) / kBlock), dim3(kBlock), 0, s, xa, xp,
                       n, rec_stride, rec_off, tgt);
    return hipGetLastError();
}

// One block: reduce the XSPH pass's block maxima and block sums, record
// e = max(0, max C) in hist[step % cap] and mean max(0, C) in
// hist[cap + step % cap], decide the rerun.  learned: 0 a plain step (record
// only), 1 a learned attempt (record, set *flag when e > bound), 2 after the
// gated rerun (when *flag is set, the rerun's e replaces the step's entry).
__global__ __launch_bounds__(kBlock) void k_pbf_decide(const int* __restrict__ maxd, int nblk, int n,
                                                       float bound, int learned, int* __restrict__ flag,
                                                       int* __restrict__ counters, float* __restrict__ hist,
                                                       int cap) {
    __shared__ int wmax[kBlock / 64];
    __shared__ float wsum[kBlock / 64];
    int m = INT_MIN;
    float s = 0.0f;
    for (int b = threadIdx.x; b < nblk; b += kBlock) {
        m = maxd[b] > m ? maxd[b] : m;
        s += __int_as_float(maxd[nblk + b]);
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        int o = __shfl_xor(m, d, 64);
        m = o > m ? o : m;
        s += __shfl_xor(s, d, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        wmax[threadIdx.x >> 6] = m;
        wsum[threadIdx.x >> 6] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        if (learned == 2 && *flag == 0) return;  // no rerun: the attempt's entry stands
        int b = wmax[0];
        float tot = wsum[0];
#pragma unroll
        for (int w = 1; w < kBlock / 64; ++w) {
            b = wmax[w] > b ? wmax[w] : b;
            tot += wsum[w];
        }
        const int bits = b >= 0 ? b : (b ^ 0x7FFFFFFF);
        float c = __int_as_float(bits);
        const float e = c > 0.0f ? c : (c == c ? 0.0f : c);  // max(0, C), NaN kept
        if (learned == 2) {
            // the rerun's final state replaces the entry of this step
            const unsigned last = (unsigned)counters[0] - 1u;
            hist[last % (unsigned)cap] = e;
            hist[cap + last % (unsigned)cap] = tot / (float)n;
            return;
        }
        const int fb = (learned && !(e <= bound)) ? 1 : 0;
        const unsigned step = (unsigned)counters[0];
        hist[step % (unsigned)cap] = e;
        hist[cap + step % (unsigned)cap] = tot / (float)n;
        counters[0] = (int)(step + 1u);
        counters[1] = counters[1] + fb;
        *flag = fb;
    }
}

hipError_t launch_pbf_decide(const int* maxd, int nblk, int n, float bound, int learned, int* flag,
                             int* counters, float* hist, int cap, hipStream_t s) {
    launch_kernel(k_pbf_decide, dim3(1), dim3(kBlock), 0, s, maxd, nblk, n, bound, learned, flag,
                       counters, hist, cap);
    return hipGetLastError();
}

// Particles of the last list-building pass whose list could not be stored
// (header count kNbrRetest): they re-test their candidates in every pass.
__global__ __launch_bounds__(kBlock) void k_pbf_count_retest(const int4* __restrict__ hdr, int n,
                                                              int* __restrict__ out) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    int c = (i < n && (hdr[i].w & 0xFFFF) == kNbrRetest) ? 1 : 0;
    for (int d = 32; d >= 1; d >>= 1) c += __shfl_xor(c, d, 64);
    if ((threadIdx.x & 63) == 0 && c) atomicAdd(out, c);
}

hipError_t launch_pbf_count_retest(const unsigned long long* nb, int n, int* out, hipStream_t s) {
    launch_kernel(k_pbf_count_retest, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s,
                  reinterpret_cast<const int4*>(nb), n, out);
    return hipGetLastError();
}

}  // namespace lus
