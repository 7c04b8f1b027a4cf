/*
 * lus_oracle.c -- CPU oracle for the Lustrine particle substep.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for tests/, __graft_entry__.smoke()
 * and the cpu_baseline leg of bench.py.  The product never links or calls it.
 *
 * Built with -O2 -fno-fast-math -ffp-contract=off (oracle/Makefile): every
 * float expression below is evaluated as written, one IEEE-754 rounding per
 * operation, no fused multiply-add.  Where the reference promotes to double
 * (std::pow(float,int), double literals) the restatement promotes the same way.
 *
 * glm 0.9.9.8 semantics used by the reference (third-party, vendored in the
 * reference tree, read as source):
 *   dot(a,b)      = (a.x*b.x + a.y*b.y) + a.z*b.z     func_geometric.inl:48-55
 *   length(v)     = sqrt(dot(v,v))                    func_geometric.inl:8-14
 *   normalize(v)  = v * (1 / sqrt(dot(v,v)))          func_geometric.inl:82-90,
 *                                                     func_exponential.inl:136-139
 *   v / s         = (v.x/s, v.y/s, v.z/s)             type_vec3.inl:715-721
 *   clamp(x,a,b)  = min(max(x,a),b) per component     func_common.inl:240-246
 */
#include "lus_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* glm-shaped helpers (float)                                          */
/* ------------------------------------------------------------------ */
static inline float v3dot(const float* a, const float* b) {
    float t0 = a[0] * b[0];
    float t1 = a[1] * b[1];
    float t2 = a[2] * b[2];
    return t0 + t1 + t2;
}
static inline float v3len(const float* a) { return sqrtf(v3dot(a, a)); }
static inline float gmin(float x, float y) { return (y < x) ? y : x; }
static inline float gmax(float x, float y) { return (x < y) ? y : x; }
static inline float gclamp(float x, float lo, float hi) { return gmin(gmax(x, lo), hi); }

/* ------------------------------------------------------------------ */
/* constants: src/Lustrine.cpp:383-386, 543-558; src/Simulation.hpp:147-170 */
/* ------------------------------------------------------------------ */
void orc_init_constants(OrcSim* s, int X, int Y, int Z, float particleRadius,
                        float particleDiameter, float kernel_radius_scale) {
    const double pi = 3.14159265358979323846; /* src/Lustrine.cpp:16 */
    s->kernelFactor = 0.5f;
    s->rest_density = 24.0f;
    s->mass = 5.0f;
    s->gravity[0] = 0.0f; s->gravity[1] = -10.0f; s->gravity[2] = 0.0f;
    s->relaxation_epsilon = 10.0f;
    s->s_corr_dq = 0.5f; s->s_corr_k = 1.0f; s->s_corr_n = 4.0f;
    s->c_xsph = 0.1f;
    s->domainX = (float)X; s->domainY = (float)Y; s->domainZ = (float)Z;
    s->particleRadius = particleRadius;
    s->particleDiameter = particleDiameter;
    s->kernelRadius = kernel_radius_scale * particleRadius;
    s->cell_size = 1.0f * s->kernelRadius;
    /* float h3 = std::pow(kernelRadius, 3): pow(double, 3) rounded to float */
    float h3 = (float)pow((double)s->kernelRadius, 3.0);
    s->cubic_kernel_k = (float)(8.0f / (pi * (double)h3));
    s->cubic_kernel_l = (float)(48.0f / (pi * (double)h3));
    s->gridX = (int)(s->domainX / s->cell_size) + 1;
    s->gridY = (int)(s->domainY / s->cell_size) + 1;
    s->gridZ = (int)(s->domainZ / s->cell_size) + 1;
    s->num_grid_cells = s->gridX * s->gridY * s->gridZ;
}

/* ------------------------------------------------------------------ */
/* SPH kernels: src/Kernels.cpp                                        */
/* ------------------------------------------------------------------ */
float orc_cubic_kernel(const OrcSim* s, float r) { /* Kernels.cpp:6-19 */
    float q = (r * s->kernelFactor) / s->kernelRadius;
    float result = 0.0f;
    if ((double)q <= 1.0) {
        if ((double)q <= 0.5) {
            float q2 = q * q;
            float q3 = q2 * q;
            float a = 6.0f * q3;
            float b = 6.0f * q2;
            result = s->cubic_kernel_k * ((a - b) + 1.0f);
        } else {
            result = s->cubic_kernel_k * (2.0f * powf(1.0f - q, 3.0f));
        }
    }
    return result;
}

void orc_cubic_kernel_grad(const OrcSim* s, const float r[3], float out[3]) { /* Kernels.cpp:26-41 */
    out[0] = out[1] = out[2] = 0.0f;
    float rl = v3len(r) * s->kernelFactor;
    float q = rl / s->kernelRadius;
    if ((double)rl > 1.0e-5 && (double)q <= 1.0) {
        float c = 1.0f / (rl * s->kernelRadius);
        float gq[3] = {c * r[0], c * r[1], c * r[2]};
        float sc;
        if ((double)q <= 0.5) {
            sc = (s->cubic_kernel_l * q) * (3.0f * q - 2.0f);
        } else {
            const float f = 1.0f - q;
            sc = s->cubic_kernel_l * ((-f) * f);
        }
        out[0] = sc * gq[0]; out[1] = sc * gq[1]; out[2] = sc * gq[2];
    }
}

float orc_poly6_kernel(const OrcSim* s, float r) { /* Kernels.cpp:43-51 */
    float result = 0.0f;
    float hf = s->kernelRadius * s->kernelFactor;
    if (r <= s->kernelRadius) {
        /* 64.0f * 3.14f is a float product; std::pow(float,int) is double. */
        float c64 = 64.0f * 3.14f;
        double c = 315.0 / ((double)c64 * pow((double)hf, 9.0));
        float fr = s->kernelFactor * r;
        double inner = pow((double)hf, 2.0) - pow((double)fr, 2.0);
        result = (float)(c * pow(inner, 3.0));
    }
    return result;
}

float orc_poly6_kernel_vec(const OrcSim* s, const float r[3]) { /* Kernels.cpp:53-55 */
    (void)r; /* glm's r.length() is the component count, 3 */
    return orc_poly6_kernel(s, 3.0f);
}

void orc_spiky_kernel(const OrcSim* s, const float r[3], float out[3]) { /* Kernels.cpp:57-67 */
    out[0] = out[1] = out[2] = 0.0f;
    float rl = v3len(r);
    if ((double)rl > 0.0 && rl <= s->kernelRadius) {
        float hf = s->kernelRadius * s->kernelFactor;
        double a = 15.0 / ((double)3.14f * pow((double)hf, 6.0));
        float b = hf - (rl * s->kernelFactor);
        float temp = (float)(a * pow((double)b, 2.0));
        float den = rl * s->kernelFactor;
        out[0] = (r[0] / den) * temp;
        out[1] = (r[1] / den) * temp;
        out[2] = (r[2] / den) * temp;
    }
}

/* ------------------------------------------------------------------ */
/* grid: src/neighbors/Utils.hpp:24-33, Sorting.cpp:11-33              */
/* ------------------------------------------------------------------ */
static inline int cell_id_of(const OrcSim* s, const float* p) {
    float px = p[0] / s->cell_size;
    float py = p[1] / s->cell_size;
    float pz = p[2] / s->cell_size;
    return ((int)py) * s->gridX * s->gridZ + ((int)px) * s->gridZ + ((int)pz);
}

void orc_cell_ids(const OrcSim* s, const float* pos, int n, int* out) {
    for (int i = 0; i < n; i++) out[i] = cell_id_of(s, pos + 3 * (size_t)i);
}

void orc_counting_sort(int* counts, const int* keys, size_t n, size_t num_cells, int* sorted) {
    memset(counts, 0, (num_cells + 1) * sizeof(int));
    for (size_t i = 0; i < n; i++) counts[keys[i]]++;
    for (size_t i = 1; i <= num_cells; i++) counts[i] += counts[i - 1];
    for (long i = (long)n - 1; i >= 0; i--) {
        sorted[counts[keys[i]] - 1] = (int)i;
        counts[keys[i]]--;
    }
}

/* growable int list used to restate std::vector<int>::push_back order */
typedef struct { int* a; int n, cap; } IList;
static void il_push(IList* l, int v) {
    if (l->n == l->cap) {
        l->cap = l->cap ? 2 * l->cap : 16;
        l->a = (int*)realloc(l->a, sizeof(int) * (size_t)l->cap);
    }
    l->a[l->n++] = v;
}

/* ------------------------------------------------------------------ */
/* fluid neighbor lists: src/neighbors/Neighbors.cpp:366-450           */
/* ------------------------------------------------------------------ */
long orc_neighbors_fluid(const OrcSim* s, const float* ps,
                         int sand_start, int sand_end, int solid_start, int solid_end,
                         long* off, int* idx, long cap) {
    const int C = s->num_grid_cells;
    /* cells: sand pushed first (ascending), then solids (ascending) :374-384 */
    IList* cells = (IList*)calloc((size_t)C, sizeof(IList));
    for (int i = sand_start; i < sand_end; i++) il_push(&cells[cell_id_of(s, ps + 3 * (size_t)i)], i);
    for (int i = solid_start; i < solid_end; i++) il_push(&cells[cell_id_of(s, ps + 3 * (size_t)i)], i);
    const float h2 = s->kernelRadius * s->kernelRadius;
    const int nsand = sand_end - sand_start;
    /* Per particle, entries appear in neighbor-cell (dy,dx,dz) order, then cell
     * order (:401-438); the sweep over current cells :386-388 does not change
     * the order inside one particle's list. */
    long total = 0;
    for (int pass = 0; pass < 2; pass++) {
        total = 0;
        for (int k = 0; k < nsand; k++) {
            int i = sand_start + k;
            if (pass == 1) off[k] = total;
            const float* self = ps + 3 * (size_t)i;
            int cid = cell_id_of(s, self);
            int yy = cid / (s->gridX * s->gridZ);
            int rem = cid - yy * s->gridX * s->gridZ;
            int xx = rem / s->gridZ, zz = rem % s->gridZ;
            for (int y = -1; y <= 1; y++)
                for (int x = -1; x <= 1; x++)
                    for (int z = -1; z <= 1; z++) {
                        if (xx + x < 0 || xx + x >= s->gridX || yy + y < 0 || yy + y >= s->gridY ||
                            zz + z < 0 || zz + z >= s->gridZ)
                            continue;
                        int nc = (yy + y) * s->gridX * s->gridZ + (xx + x) * s->gridZ + (zz + z);
                        const IList* l = &cells[nc];
                        for (int t = 0; t < l->n; t++) {
                            int j = l->a[t];
                            const float* o = ps + 3 * (size_t)j;
                            float d[3] = {self[0] - o[0], self[1] - o[1], self[2] - o[2]};
                            if (v3dot(d, d) <= h2) {
                                if (pass == 1) idx[total] = j;
                                total++;
                            }
                        }
                    }
        }
        if (pass == 0 && total > cap) { total = -total; break; }
        if (pass == 1) off[nsand] = total;
    }
    for (int c = 0; c < C; c++) free(cells[c].a);
    free(cells);
    return total;
}

/* ------------------------------------------------------------------ */
/* sand path: find_neighbors_uniform_grid_v1 (Neighbors.cpp:262-364)    */
/* and simulate_sand (Simulate.cpp:156-325)                            */
/* ------------------------------------------------------------------ */
static void sand_neighbors_v1(const OrcSim* s, OrcSandState* st, IList* nbrs) {
    const int C = s->num_grid_cells;
    const int n = st->num_sand_particles;
    float* ps = st->positions_star;
    /* clear_neighbors (:264 -> :21-25) is the caller's fresh nbrs[] */
    if (!st->computed_static_particles) { /* :266-272 */
        int* cnt = (int*)calloc((size_t)C + 1, sizeof(int));
        int nsol = st->ptr_solid_end - st->ptr_solid_start;
        int* cid = (int*)malloc(sizeof(int) * (size_t)(nsol > 0 ? nsol : 1));
        for (int i = st->ptr_solid_start; i < st->ptr_solid_end; i++) {
            cid[i - st->ptr_solid_start] = cell_id_of(s, ps + 3 * (size_t)i);
            cnt[cid[i - st->ptr_solid_start] + 1]++;
        }
        free(st->static_off); free(st->static_idx);
        st->static_off = (long*)malloc(sizeof(long) * ((size_t)C + 1));
        st->static_off[0] = 0;
        for (int c = 0; c < C; c++) st->static_off[c + 1] = st->static_off[c] + cnt[c + 1];
        st->static_idx = (int*)malloc(sizeof(int) * (size_t)(nsol > 0 ? nsol : 1));
        long* fill = (long*)malloc(sizeof(long) * (size_t)C);
        memcpy(fill, st->static_off, sizeof(long) * (size_t)C);
        for (int i = st->ptr_solid_start; i < st->ptr_solid_end; i++)
            st->static_idx[fill[cid[i - st->ptr_solid_start]]++] = i;
        free(fill); free(cid); free(cnt);
        st->computed_static_particles = 1;
    }
    int* keys = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    int* sorted = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    int* counts = (int*)malloc(sizeof(int) * ((size_t)C + 1));
    for (int i = st->ptr_sand_start; i < st->ptr_sand_end; i++)
        keys[i] = cell_id_of(s, ps + 3 * (size_t)i); /* :275-278 */
    orc_counting_sort(counts, keys, (size_t)n, (size_t)C, sorted); /* :280-286 */
    free(st->sorted_indices);
    st->sorted_indices = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    memcpy(st->sorted_indices, sorted, sizeof(int) * (size_t)n);
    /* gather into sorted order (:288-304) */
    float* tp = (float*)malloc(sizeof(float) * 3 * (size_t)(n > 0 ? n : 1));
    float* tps = (float*)malloc(sizeof(float) * 3 * (size_t)(n > 0 ? n : 1));
    float* tv = (float*)malloc(sizeof(float) * 3 * (size_t)(n > 0 ? n : 1));
    int* ta = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    memcpy(tp, st->positions, sizeof(float) * 3 * (size_t)n);
    memcpy(tps, ps, sizeof(float) * 3 * (size_t)n);
    memcpy(tv, st->velocities, sizeof(float) * 3 * (size_t)n);
    memcpy(ta, st->attracted, sizeof(int) * (size_t)n);
    /* sand part of every cell, in slot order: CSR sand_start/sand_cnt */
    long* sand_off = (long*)calloc((size_t)C + 1, sizeof(long));
    int* cell_of = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    for (int i = st->ptr_sand_start; i < st->ptr_sand_end; i++) {
        int o = sorted[i];
        memcpy(st->positions + 3 * (size_t)i, tp + 3 * (size_t)o, sizeof(float) * 3);
        memcpy(ps + 3 * (size_t)i, tps + 3 * (size_t)o, sizeof(float) * 3);
        memcpy(st->velocities + 3 * (size_t)i, tv + 3 * (size_t)o, sizeof(float) * 3);
        st->attracted[i] = ta[o];
        cell_of[i] = cell_id_of(s, ps + 3 * (size_t)i);
        sand_off[cell_of[i] + 1]++;
    }
    for (int c = 0; c < C; c++) sand_off[c + 1] += sand_off[c];
    /* slots are cell-sorted, so cell c's sand entries are slots
     * [sand_off[c], sand_off[c+1]) in ascending order */
    const float h2 = s->kernelRadius * s->kernelRadius;
    for (int i = st->ptr_sand_start; i < st->ptr_sand_end; i++) { /* :306-361 */
        int cid = cell_of[i];
        int yy = cid / (s->gridX * s->gridZ);
        int tmp = cid - yy * s->gridX * s->gridZ;
        int zz = tmp % s->gridZ;
        int xx = tmp / s->gridZ;
        int yl = -1, yu = 1, xl = -1, xu = 1, zl = -1, zu = 1;
        if (yy + yl < 0) yl = 0;
        if (yy + yu >= s->gridY) yu = 0;
        if (xx + xl < 0) xl = 0;
        if (xx + xu >= s->gridX) xu = 0;
        if (zz + zl < 0) zl = 0;
        if (zz + zu >= s->gridZ) zu = 0;
        const float* self = ps + 3 * (size_t)i;
        for (int y = yl; y <= yu; y++)
            for (int x = xl; x <= xu; x++)
                for (int z = zl; z <= zu; z++) {
                    int nc = (yy + y) * s->gridX * s->gridZ + (xx + x) * s->gridZ + (zz + z);
                    /* cell = static solids first, then sand (:293, :303) */
                    long ns0 = st->static_off[nc], ns1 = st->static_off[nc + 1];
                    long nd0 = sand_off[nc], nd1 = sand_off[nc + 1];
                    long cnt = (ns1 - ns0) + (nd1 - nd0);
                    for (long t = 0; t < cnt; t++) {
                        int j = (t < ns1 - ns0) ? st->static_idx[ns0 + t] : (int)(nd0 + (t - (ns1 - ns0)));
                        if (j < i) continue;
                        const float* o = ps + 3 * (size_t)j;
                        float d[3] = {self[0] - o[0], self[1] - o[1], self[2] - o[2]};
                        if (v3dot(d, d) <= h2) {
                            il_push(&nbrs[i], j);
                            if (j < st->ptr_solid_start) il_push(&nbrs[j], i);
                        }
                    }
                }
    }
    free(keys); free(sorted); free(counts); free(tp); free(tps); free(tv); free(ta);
    free(cell_of);
    /* the cell lists stay as they are until the next search: the sinks and
     * the cell queries of Lustrine.cpp read them after the step */
    free(st->cell_sand_off);
    st->cell_sand_off = sand_off;
}

/* get_cell_indices (src/neighbors/Utils.hpp:13-21): clamp to
 * [cs/2, domain - cs/2] per axis, divide by the cell size, truncate */
static void cell_indices(const OrcSim* s, const float* pos, int* out) {
    const float half = s->cell_size * 0.5f;
    const float hi[3] = {s->domainX - half, s->domainY - half, s->domainZ - half};
    for (int a = 0; a < 3; a++) {
        float q = pos[a] < half ? half : pos[a];
        q = q > hi[a] ? hi[a] : q;
        q = q / s->cell_size;
        out[a] = (int)q;
    }
}

int orc_query_cell_num_particles(const OrcSim* s, const OrcSandState* st, const float* min_pos,
                                 const float* max_pos, int include_solid) {
    if (!st->cell_sand_off) return 0;
    int lo[3], hi[3];
    cell_indices(s, min_pos, lo);
    cell_indices(s, max_pos, hi);
    int counter = 0;
    for (int y = lo[1]; y <= hi[1]; y++)
        for (int x = lo[0]; x <= hi[0]; x++)
            for (int z = lo[2]; z <= hi[2]; z++) {
                const int c = y * s->gridX * s->gridZ + x * s->gridZ + z;
                const long s0 = st->static_off[c], s1 = st->static_off[c + 1];
                const long d0 = st->cell_sand_off[c], d1 = st->cell_sand_off[c + 1];
                if (include_solid) {
                    counter += (int)((s1 - s0) + (d1 - d0));
                } else {
                    for (long t = s0; t < s1; t++) counter += st->static_idx[t] < st->ptr_sand_end;
                    for (long t = d0; t < d1; t++) counter += t < st->ptr_sand_end;
                }
            }
    return counter;
}

int orc_sink_evict(const OrcSim* s, OrcSandState* st, const int* cells, int ncells) {
    if (!st->cell_sand_off) return 0;
    /* temp_removal: the ids below ptr_sand_end of the listed cells, in list
     * order (static solids first, then sand slots) */
    long m = 0, cap = 16;
    int* rem = (int*)malloc(sizeof(int) * (size_t)cap);
    for (int k = 0; k < ncells; k++) {
        const int c = cells[k];
        const long s0 = st->static_off[c], s1 = st->static_off[c + 1];
        const long d0 = st->cell_sand_off[c], d1 = st->cell_sand_off[c + 1];
        for (long t = 0; t < (s1 - s0) + (d1 - d0); t++) {
            const int id = t < s1 - s0 ? st->static_idx[s0 + t] : (int)(d0 + (t - (s1 - s0)));
            if (id >= st->ptr_sand_end) continue;
            if (m == cap) {
                cap *= 2;
                rem = (int*)realloc(rem, sizeof(int) * (size_t)cap);
            }
            rem[m++] = id;
        }
    }
    /* swap-remove in turn with the last sand slot; the reference would read
     * slot -1 once the sand range is empty, so the removal stops there */
    int removed = 0;
    for (long t = 0; t < m && st->ptr_sand_end > 0; t++) {
        const int e = rem[t], last = st->ptr_sand_end - 1;
        memmove(st->positions + 3 * (size_t)e, st->positions + 3 * (size_t)last, sizeof(float) * 3);
        memmove(st->positions_star + 3 * (size_t)e, st->positions_star + 3 * (size_t)last, sizeof(float) * 3);
        memmove(st->velocities + 3 * (size_t)e, st->velocities + 3 * (size_t)last, sizeof(float) * 3);
        st->attracted[e] = st->attracted[last];
        st->ptr_sand_end--;
        st->num_sand_particles--;
        removed++;
    }
    free(rem);
    (void)s;
    return removed;
}

/* solve_boundary_collision_constraint (src/Simulate.cpp:119-131) */
static void boundary_constraint(const float n[3], const float p0[3], const float* p, float d, float out[3]) {
    float rel[3] = {p[0] - p0[0], p[1] - p0[1], p[2] - p0[2]};
    float C = v3dot(n, rel) - d;
    out[0] = out[1] = out[2] = 0.0f;
    if (C >= 0) return;
    float sv = C / v3dot(n, n);
    for (int c = 0; c < 3; c++) out[c] = (-sv) * n[c];
}

/* variant 0: simulate_sand (src/Simulate.cpp:156-325); 1: simulate_sand_credits
 * (:327-510); 2: simulate_sand_v3 (:513-667).  The three share the neighbour
 * search and the contact + friction projection; they differ in the locals
 * (:159-163, :330-334, :516-520), the integration loop, and what follows the
 * contacts (box clamp per substep for 0 and 1; for 2, four identical passes
 * from the same positions_tmp, then the six half-space boundary projections
 * from positions_tmp, :601-616). */
int orc_simulate_sand_variant(const OrcSim* s, OrcSandState* st, float dt, int variant) {
    const float collision_coeff = 0.8f;
    const float boundary_collision_coeff = 0.9f;
    const float friction_coeff = variant == 2 ? 0.5f : 0.7f;
    const float mu_s = variant == 0 ? 0.95f : 0.8f;
    const float mu_k = variant == 0 ? 0.9f : (variant == 1 ? 0.7f : 0.8f);
    if (variant == 2) dt = gclamp(dt, 0.001f, 0.016f); /* :522 */
    const float time_step = dt;
    const float X = s->domainX, Y = s->domainY, Z = s->domainZ;
    const float r = s->particleRadius;
    const float hiv[3] = {X - r, Y - r, Z - r};
    float* p = st->positions;
    float* ps = st->positions_star;
    float* v = st->velocities;
    /* integration :188-213 */
    for (int i = st->ptr_sand_start; i < st->ptr_sand_end; i++) {
        float w = 1.0f / s->mass;
        float* vi = v + 3 * (size_t)i;
        float* pi = p + 3 * (size_t)i;
        if (variant == 2) { /* :541-556 */
            for (int c = 0; c < 3; c++) vi[c] = vi[c] + s->gravity[c] * dt;
            if (st->attract_flag) {
                float d[3] = {st->player_position[0] - pi[0], st->player_position[1] - pi[1],
                              st->player_position[2] - pi[2]};
                float inv = 1.0f / sqrtf(v3dot(d, d));
                for (int c = 0; c < 3; c++) {
                    float t = d[c] * inv;
                    t = t * r;
                    t = t * 1000.0f;
                    t = t * dt;
                    t = t * w;
                    vi[c] = vi[c] + t;
                }
            }
            if (st->blow_flag) {
                float d[3] = {st->player_position[0] - pi[0], st->player_position[1] - pi[1],
                              st->player_position[2] - pi[2]};
                float inv = 1.0f / sqrtf(v3dot(d, d));
                for (int c = 0; c < 3; c++) {
                    float t = -(d[c] * inv);
                    t = t * r;
                    t = t * 200.0f;
                    t = t * w;
                    vi[c] = vi[c] + t;
                }
            }
            float* psi = ps + 3 * (size_t)i;
            for (int c = 0; c < 3; c++) psi[c] = pi[c] + vi[c] * dt; /* no clamp */
            continue;
        }
        if (variant == 1 && (st->attracted[i] & 2)) { /* :361-364 credits fall at 1 */
            vi[0] = 0.0f;
            vi[1] = -1.0f;
            vi[2] = 0.0f;
        }
        for (int c = 0; c < 3; c++) vi[c] = vi[c] + s->gravity[c] * dt;
        if (st->prev_attract_flag && !st->attract_flag) {
            if (variant == 1) st->attracted[i] &= ~1;
            else st->attracted[i] = 0;
        }
        if (st->attract_flag) {
            float d0[3] = {st->player_position[0] - pi[0], st->player_position[1] - pi[1],
                           st->player_position[2] - pi[2]};
            if (v3len(d0) < st->attract_radius) {
                if (variant == 1) {
                    st->attracted[i] |= 1;
                    st->attracted[i] &= ~2;
                } else {
                    st->attracted[i] = 1;
                }
            }
            if (variant == 1 ? (st->attracted[i] & 1) : st->attracted[i]) {
                float o[3] = {st->player_position[0] + 0.0f, st->player_position[1] + 1.5f,
                              st->player_position[2] + 0.0f};
                float d[3] = {o[0] - pi[0], o[1] - pi[1], o[2] - pi[2]};
                float inv = 1.0f / sqrtf(v3dot(d, d));
                for (int c = 0; c < 3; c++) {
                    float t = d[c] * inv;
                    t = t * 1.0f; /* attract_kernel(x) == 1 (:138-140) */
                    t = t * st->attract_coeff;
                    t = t * r;
                    t = t * dt;
                    t = t * w;
                    vi[c] = vi[c] + t;
                }
            }
        }
        if (st->blow_flag) {
            float d[3] = {st->player_position[0] - pi[0], st->player_position[1] - pi[1],
                          st->player_position[2] - pi[2]};
            float l = v3len(d);
            if (l < st->blow_radius) {
                float inv = 1.0f / sqrtf(v3dot(d, d));
                float k = (l > st->blow_radius) ? 0.0f : (1.0f - l / st->blow_radius); /* :143-154 */
                for (int c = 0; c < 3; c++) {
                    float t = -(d[c] * inv);
                    t = t * k;
                    t = t * st->blow_coeff;
                    t = t * r;
                    t = t * w;
                    vi[c] = vi[c] + t;
                }
                if (variant == 1) st->attracted[i] &= ~2;
            }
        }
        float* psi = ps + 3 * (size_t)i;
        for (int c = 0; c < 3; c++) {
            psi[c] = pi[c] + vi[c] * dt;
            psi[c] = gclamp(psi[c], r, hiv[c]);
        }
    }
    /* neighbors :215 */
    const int nslots = st->ptr_sand_end > 0 ? st->ptr_sand_end : 1;
    IList* nbrs = (IList*)calloc((size_t)nslots, sizeof(IList));
    sand_neighbors_v1(s, st, nbrs);
    p = st->positions; ps = st->positions_star; v = st->velocities;
    /* positions_tmp :217-223 (variant 2: every slot on every call, :562) */
    if (variant == 2) {
        memcpy(st->positions_tmp, ps, sizeof(float) * 3 * st->total_allocated);
    } else if (st->first_iteration) {
        memcpy(st->positions_tmp, ps, sizeof(float) * 3 * st->total_allocated);
        st->first_iteration = 0;
    } else {
        memcpy(st->positions_tmp, ps, sizeof(float) * 3 * (size_t)st->num_sand_particles);
    }
    float* pt = st->positions_tmp;
    const float dd = s->particleDiameter;
    /* variants 0 and 1: 4 Jacobi substeps (:226-311); variant 2: the 4
     * substeps sit inside the particle loop and all read the same
     * positions_tmp (:566-633) */
    const int outer = variant == 2 ? 1 : 4, inner = variant == 2 ? 4 : 1;
    for (int substep = 0; substep < outer; ++substep) {
        for (int i = st->ptr_sand_start; i < st->ptr_sand_end; i++) {
          for (int rep = 0; rep < inner; ++rep) {
            float dp[3] = {0.0f, 0.0f, 0.0f};
            for (int t = 0; t < nbrs[i].n; t++) {
                int j = nbrs[i].a[t];
                if (i == j) continue;
                const float* pi = pt + 3 * (size_t)i;
                const float* pj = pt + 3 * (size_t)j;
                float ij[3] = {pi[0] - pj[0], pi[1] - pj[1], pi[2] - pj[2]};
                /* zero-length fix of :236-239 (absent from variant 2) */
                if (variant != 2 && v3len(ij) == 0.0f) { ij[0] = 0.0f; ij[1] = 0.00001f; ij[2] = 0.0f; }
                float len = v3len(ij);
                if (len > dd) continue;
                float tmp[3], xid[3], xjd[3], nx[3], rel[3], xt[3];
                if (j >= st->ptr_sand_start && j < st->ptr_sand_end) {
                    float mass = s->mass, nmass = s->mass;
                    float sc = collision_coeff * nmass;
                    sc = sc / (mass + nmass);
                    sc = sc * (len - dd);
                    for (int c = 0; c < 3; c++) tmp[c] = (sc * ij[c]) / len;
                    for (int c = 0; c < 3; c++) dp[c] = dp[c] - tmp[c];
                    float d = v3len(tmp);
                    const float* pos_i = p + 3 * (size_t)i;
                    const float* pos_j = p + 3 * (size_t)j;
                    for (int c = 0; c < 3; c++) {
                        xid[c] = (pi[c] - tmp[c]) - pos_i[c];
                        xjd[c] = (pj[c] + tmp[c]) - pos_j[c];
                        nx[c] = (pi[c] - tmp[c]) - (pj[c] + tmp[c]);
                    }
                    float inv = 1.0f / sqrtf(v3dot(nx, nx));
                    for (int c = 0; c < 3; c++) nx[c] = nx[c] * inv;
                    for (int c = 0; c < 3; c++) rel[c] = xid[c] - xjd[c];
                    float rd = v3dot(rel, nx);
                    for (int c = 0; c < 3; c++) xt[c] = rel[c] - rd * nx[c];
                    if ((d * mu_s) > (v3len(xt) + 1e-9f)) {
                        for (int c = 0; c < 3; c++) dp[c] = dp[c] - friction_coeff * xt[c];
                    } else {
                        float q = (mu_k * d) / (v3len(xt) + 1e-9f);
                        float ratio = q < 1 ? q : 1.0f;
                        for (int c = 0; c < 3; c++) dp[c] = dp[c] - (friction_coeff * xt[c]) * ratio;
                    }
                } else {
                    float sc = collision_coeff * (len - dd);
                    for (int c = 0; c < 3; c++) tmp[c] = (sc * ij[c]) / len;
                    for (int c = 0; c < 3; c++) dp[c] = dp[c] - tmp[c];
                    float d = v3len(tmp);
                    const float* pos_i = p + 3 * (size_t)i;
                    for (int c = 0; c < 3; c++) {
                        xid[c] = (pi[c] - tmp[c]) - pos_i[c];
                        xjd[c] = 0.0f;
                        nx[c] = (pi[c] - tmp[c]) - pj[c];
                    }
                    float inv = 1.0f / sqrtf(v3dot(nx, nx));
                    for (int c = 0; c < 3; c++) nx[c] = nx[c] * inv;
                    for (int c = 0; c < 3; c++) rel[c] = xid[c] - xjd[c];
                    float rd = v3dot(rel, nx);
                    for (int c = 0; c < 3; c++) xt[c] = rel[c] - rd * nx[c];
                    if ((d * mu_s) > (v3len(xt) + 1e-9f)) {
                        for (int c = 0; c < 3; c++) dp[c] = dp[c] - friction_coeff * xt[c];
                    } else {
                        float q = (mu_k * d) / (v3len(xt) + 1e-9f);
                        float ratio = q < 1 ? q : 1.0f;
                        for (int c = 0; c < 3; c++) dp[c] = dp[c] - (friction_coeff * xt[c]) * ratio;
                    }
                }
                if (variant == 1 && (st->attracted[i] & 2)) { /* :463-469 */
                    /* the three draws are made and not used */
                    (void)rand();
                    (void)rand();
                    (void)rand();
                    st->attracted[i] &= ~2;
                }
            }
            for (int c = 0; c < 3; c++) ps[3 * (size_t)i + c] = pt[3 * (size_t)i + c] + dp[c];
          }
        }
        if (variant == 2) { /* :601-616: six half-spaces, evaluated at positions_tmp */
            const float wn[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
            const float w0[6][3] = {{0, 0, 0}, {X, 0, 0}, {0, 0, 0}, {0, Y, 0}, {0, 0, 0}, {0, 0, Z}};
            for (int i = st->ptr_sand_start; i < st->ptr_sand_end; i++) {
                const float* q = pt + 3 * (size_t)i;
                float dp[3] = {0.0f, 0.0f, 0.0f};
                for (int wall = 0; wall < 6; wall++) {
                    float bc[3];
                    boundary_constraint(wn[wall], w0[wall], q, r, bc);
                    for (int c = 0; c < 3; c++) dp[c] = dp[c] + boundary_collision_coeff * bc[c];
                }
                for (int c = 0; c < 3; c++) ps[3 * (size_t)i + c] = ps[3 * (size_t)i + c] + dp[c];
            }
        } else {
            for (int i = st->ptr_sand_start; i < st->ptr_sand_end; i++) /* :296-308 */
                for (int c = 0; c < 3; c++) ps[3 * (size_t)i + c] = gclamp(ps[3 * (size_t)i + c], r, hiv[c]);
        }
        memcpy(pt + 3 * (size_t)st->ptr_sand_start, ps + 3 * (size_t)st->ptr_sand_start,
               sizeof(float) * 3 * (size_t)st->num_sand_particles); /* :310, :619 */
    }
    for (int i = st->ptr_sand_start; i < st->ptr_sand_end; i++) /* :316-319 */
        for (int c = 0; c < 3; c++) {
            v[3 * (size_t)i + c] = (ps[3 * (size_t)i + c] - p[3 * (size_t)i + c]) / time_step;
            p[3 * (size_t)i + c] = ps[3 * (size_t)i + c];
        }
    if (variant != 2) st->prev_attract_flag = st->attract_flag; /* :323, :508 */
    /* flatten the neighbor lists for inspection */
    long total = 0;
    for (int i = st->ptr_sand_start; i < st->ptr_sand_end; i++) total += nbrs[i].n;
    if (total > st->nbr_cap) {
        free(st->nbr_idx);
        st->nbr_idx = (int*)malloc(sizeof(int) * (size_t)(total > 0 ? total : 1));
        st->nbr_cap = total;
    }
    free(st->nbr_off);
    st->nbr_off = (long*)malloc(sizeof(long) * ((size_t)st->num_sand_particles + 1));
    long o = 0;
    for (int i = st->ptr_sand_start; i < st->ptr_sand_end; i++) {
        st->nbr_off[i - st->ptr_sand_start] = o;
        if (nbrs[i].n) memcpy(st->nbr_idx + o, nbrs[i].a, sizeof(int) * (size_t)nbrs[i].n);
        o += nbrs[i].n;
        free(nbrs[i].a);
    }
    st->nbr_off[st->num_sand_particles] = o;
    for (int i = st->ptr_sand_end; i < nslots; i++) free(nbrs[i].a);
    free(nbrs);
    return 0;
}

int orc_simulate_sand(const OrcSim* s, OrcSandState* st, float dt) {
    return orc_simulate_sand_variant(s, st, dt, 0);
}

void orc_sand_state_free(OrcSandState* st) {
    free(st->static_off); free(st->static_idx); free(st->sorted_indices);
    free(st->nbr_off); free(st->nbr_idx); free(st->cell_sand_off);
    st->static_off = NULL; st->static_idx = NULL; st->sorted_indices = NULL;
    st->nbr_off = NULL; st->nbr_idx = NULL; st->nbr_cap = 0; st->cell_sand_off = NULL;
}

/* ------------------------------------------------------------------ */
/* fluid path: simulate_fluid, src/Simulate.cpp:7-115                  */
/* ------------------------------------------------------------------ */
static float resolve_collision(float value, float mn, float mx) { /* :11-24 */
    const float eps = 0.01f;
    if (value <= mn) return eps;
    if (value > mx) return mx - eps;
    return value;
}

int orc_simulate_fluid(OrcSim* s, float* p, float* ps, float* v, float* lambdas,
                       int sand_start, int sand_end, int solid_start, int solid_end, float dt) {
    dt = gclamp(dt, 0.001f, 0.01f); /* :31 */
    const float time_step = dt;
    const int X = (int)s->domainX, Y = (int)s->domainY, Z = (int)s->domainZ;
    for (int i = sand_start; i < sand_end; i++) { /* :48-51 */
        for (int c = 0; c < 3; c++) {
            float g = s->gravity[c] * s->mass;
            g = g * dt;
            v[3 * (size_t)i + c] = v[3 * (size_t)i + c] + g;
        }
        for (int c = 0; c < 3; c++) ps[3 * (size_t)i + c] = p[3 * (size_t)i + c] + v[3 * (size_t)i + c] * dt;
    }
    const int nsand = sand_end - sand_start;
    long cap = (long)nsand * 64 + 64;
    long* off = (long*)malloc(sizeof(long) * ((size_t)nsand + 1));
    int* idx = (int*)malloc(sizeof(int) * (size_t)cap);
    long tot = orc_neighbors_fluid(s, ps, sand_start, sand_end, solid_start, solid_end, off, idx, cap);
    if (tot < 0) {
        cap = -tot;
        idx = (int*)realloc(idx, sizeof(int) * (size_t)cap);
        tot = orc_neighbors_fluid(s, ps, sand_start, sand_end, solid_start, solid_end, off, idx, cap);
    }
    const float W0 = orc_cubic_kernel(s, 0.0f);
    for (int i = sand_start; i < sand_end; i++) { /* :58-88 */
        long a = off[i - sand_start], b = off[i - sand_start + 1];
        float density = 0.0f;
        for (long t = a; t < b; t++) {
            int j = idx[t];
            float ij[3] = {ps[3 * (size_t)i] - ps[3 * (size_t)j], ps[3 * (size_t)i + 1] - ps[3 * (size_t)j + 1],
                           ps[3 * (size_t)i + 2] - ps[3 * (size_t)j + 2]};
            float len = v3len(ij);
            density = density + s->mass * orc_cubic_kernel(s, len);
        }
        density = density + s->mass * W0;
        float constraint_i = (float)((double)(density / s->rest_density) - 1.0);
        float gsum = 0.0f;
        float gcp[3] = {0.0f, 0.0f, 0.0f};
        for (long t = a; t < b; t++) {
            int j = idx[t];
            float tmp[3] = {ps[3 * (size_t)i] - ps[3 * (size_t)j], ps[3 * (size_t)i + 1] - ps[3 * (size_t)j + 1],
                            ps[3 * (size_t)i + 2] - ps[3 * (size_t)j + 2]};
            float g[3];
            orc_cubic_kernel_grad(s, tmp, g);
            float m = -(s->mass / s->rest_density);
            float ng[3] = {m * g[0], m * g[1], m * g[2]};
            gsum = gsum + v3dot(ng, ng);
            for (int c = 0; c < 3; c++) gcp[c] = gcp[c] - ng[c];
        }
        gsum = gsum + v3dot(gcp, gcp);
        lambdas[i] = 0.0f;
        if ((double)gsum > 0.0) lambdas[i] = -constraint_i / (gsum + s->relaxation_epsilon);
    }
    const float Wdq = orc_cubic_kernel(s, s->s_corr_dq);
    for (int i = sand_start; i < sand_end; i++) { /* :90-113 */
        long a = off[i - sand_start], b = off[i - sand_start + 1];
        float pf[3] = {0.0f, 0.0f, 0.0f};
        for (long t = a; t < b; t++) {
            int j = idx[t];
            float ij[3] = {ps[3 * (size_t)i] - ps[3 * (size_t)j], ps[3 * (size_t)i + 1] - ps[3 * (size_t)j + 1],
                           ps[3 * (size_t)i + 2] - ps[3 * (size_t)j + 2]};
            float sc = -s->s_corr_k * powf(orc_cubic_kernel(s, v3len(ij)) / Wdq, s->s_corr_n); /* :7-9 */
            float coef = (lambdas[i] + lambdas[t - a]) + sc; /* lambdas[j], j = list position */
            float g[3];
            orc_cubic_kernel_grad(s, ij, g);
            for (int c = 0; c < 3; c++) pf[c] = pf[c] + coef * g[c];
        }
        for (int c = 0; c < 3; c++) pf[c] = pf[c] / s->rest_density;
        float* q = ps + 3 * (size_t)i;
        for (int c = 0; c < 3; c++) q[c] = q[c] + pf[c];
        q[0] = resolve_collision(q[0], s->particleRadius, (float)X - s->particleRadius);
        q[1] = resolve_collision(q[1], s->particleRadius, (float)Y - s->particleRadius);
        q[2] = resolve_collision(q[2], s->particleRadius, (float)Z - s->particleRadius);
        for (int c = 0; c < 3; c++) {
            v[3 * (size_t)i + c] = (q[c] - p[3 * (size_t)i + c]) / time_step;
            p[3 * (size_t)i + c] = q[c];
        }
    }
    free(off); free(idx);
    return 0;
}

/* ------------------------------------------------------------------ */
/* position-based fluids substep, FP32 and FP64 twins                   */
/* ------------------------------------------------------------------ */
#define REAL float
#define SFX(name) name##_f32
#define RSQRT sqrtf
#include "pbf_core.inc"
#include "mixed_core.inc"
#undef REAL
#undef SFX
#undef RSQRT

#define REAL double
#define SFX(name) name##_f64
#define RSQRT sqrt
#include "pbf_core.inc"
#include "mixed_core.inc"
#undef REAL
#undef SFX
#undef RSQRT

double orc_pbf_density_error_f64(const OrcPbfParams* p, int n, const double* x) {
    return pbf_density_error_f64(p, n, x);
}
