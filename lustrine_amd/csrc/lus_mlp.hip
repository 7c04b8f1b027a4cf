// lus_mlp.hip -- the per-particle warm-start network on gfx950 matrix cores.
//
// Network: 32 inputs -> 64 (SiLU) -> 64 (SiLU) -> 3.  BF16 weights and
// activations, FP32 accumulation (v_mfma_f32_16x16x32_bf16), FP32 bias and
// SiLU.  Parameter order (FP32 master copy): W1[32][64] b1[64] W2[64][64]
// b2[64] W3[64][3] b3[3]; W[k][n] maps input k to output n, so each (W, b)
// pair is one [fan_in + 1][fan_out] matrix whose last row is the bias.
//
// Operand layout of v_mfma_f32_16x16x32_bf16 (wave64, lane l):
//   A (16 x 32, rows m):   lane l holds A[l % 16][8 * (l / 16) + e], e = 0..7
//   B (32 x 16, cols n):   lane l holds B[8 * (l / 16) + e][l % 16], e = 0..7
//   C (16 x 16):           lane l holds C[4 * (l / 16) + i][l % 16], i = 0..3
// The GPU tests compare the forward pass and the gradients with a float64 CPU
// implementation that rounds to bf16 at the same points, which fixes these
// layouts.
#include <hip/hip_runtime.h>

#include "lus_device.hpp"
#include "lus_internal.hpp"
#include "lus_launch.hpp"

namespace lus {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ float silu(float z) { return z / (1.0f + __expf(-z)); }
__device__ __forceinline__ float silu_grad(float z) {
    const float s = 1.0f / (1.0f + __expf(-z));
    return s * (1.0f + z * (1.0f - s));
}

// ------------------------------------------------------------------ inference
// Inference runs each layer transposed, Y^T = W^T X^T: the weights are the A
// operand and the activations the B operand, so the C tile of a layer holds,
// in lane l, units 16j + 4(l/16) + i (i = 0..3) of sample l%16.  The next
// layer's reduction index is permuted to take those values as they lie:
// k-step s, lane group g = l/16, element e  <->  hidden unit
//   pi(s, g, e) = 16 (2s + e/4) + 4g + e%4,
// i.e. the B operand of k-step s is C tiles 2s and 2s+1 of the lane itself,
// and no activation leaves its lane.  A-operand images of the weights (one
// 16-byte load per lane and image):
//   [0, 4)    W1 m-tile j:            lane l: W1[8(l/16)+e][16j + l%16]
//   [4, 12)   W2 k-step s, m-tile j:  lane l: W2[pi(s, l/16, e)][16j + l%16]
//   [12, 14)  W3 k-step s (m < 3):    lane l: W3[pi(s, l/16, e)][l%16], 0 for l%16 >= 3
__device__ __forceinline__ int mlp_pi(int s, int g, int e) { return 16 * (2 * s + e / 4) + 4 * g + e % 4; }

__global__ void k_mlp_pack(const float* __restrict__ p, bf16x8* __restrict__ w) {
    const int t = threadIdx.x + blockIdx.x * blockDim.x;
    const int img = t / 64, l = t % 64;
    const int col = l % 16, grp = l / 16;
    if (img >= 14) return;
    bf16x8 v;
    if (img < 4) {
        for (int e = 0; e < 8; ++e) v[e] = (__bf16)p[kMlpW1 + (8 * grp + e) * 64 + 16 * img + col];
    } else if (img < 12) {
        const int s = (img - 4) / 4, j = (img - 4) % 4;
        for (int e = 0; e < 8; ++e) v[e] = (__bf16)p[kMlpW2 + mlp_pi(s, grp, e) * 64 + 16 * j + col];
    } else {
        const int s = img - 12;
        for (int e = 0; e < 8; ++e)
            v[e] = col < 3 ? (__bf16)p[kMlpW3 + mlp_pi(s, grp, e) * 3 + col] : (__bf16)0.0f;
    }
    w[t] = v;
}

hipError_t launch_mlp_pack(const float* params, void* packed, hipStream_t s) {
    launch_kernel(k_mlp_pack, dim3(4), dim3(256), 0, s, params, (bf16x8*)packed);
    return hipGetLastError();
}

// One wave per 16-sample tile (grid-stride over tiles).  X: bf16 [n][32]; a
// lane loads 8 features of one sample, which is the B operand of layer 1 as
// it lies.  The activations stay in registers (see the pack layout above) and
// the lanes of group 0 hold the three outputs of their sample at the end.  The
// next tile's features are requested before the current tile's layers run.
constexpr int kMlpWaves = 4;

// SiLU for inference with the hardware reciprocal (1 ulp): the activation is
// rounded to bf16 right after, far coarser than that error.
__device__ __forceinline__ float silu_fast(float z) {
    return z * __builtin_amdgcn_rcpf(1.0f + __expf(-z));
}

__global__ __launch_bounds__(64 * kMlpWaves) void k_mlp_forward(
    const bf16x8* __restrict__ X, int n, const float* __restrict__ params,
    const bf16x8* __restrict__ wpk, float out_scale, float* __restrict__ out,
    const float4* __restrict__ xp, float4* __restrict__ xa, PbfK k) {
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int col = lane % 16, grp = lane / 16;
    bf16x8 W1[4], W2[2][4], W3[2];
#pragma unroll
    for (int j = 0; j < 4; ++j) W1[j] = wpk[j * 64 + lane];
#pragma unroll
    for (int s = 0; s < 2; ++s) {
#pragma unroll
        for (int j = 0; j < 4; ++j) W2[s][j] = wpk[(4 + s * 4 + j) * 64 + lane];
        W3[s] = wpk[(12 + s) * 64 + lane];
    }
    // biases of the units this lane's C values belong to: 16j + 4 grp + i
    f32x4 B1[4], B2[4];
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            B1[j][i] = params[kMlpB1 + 16 * j + 4 * grp + i];
            B2[j][i] = params[kMlpB2 + 16 * j + 4 * grp + i];
        }
    float B3[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) B3[c] = params[kMlpB3 + c];
    const f32x4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
    const int ntiles = (n + 15) / 16;
    const int nwaves = gridDim.x * kMlpWaves;
    int tile = blockIdx.x * kMlpWaves + w;  // uniform per wave: the loop below never diverges
    bf16x8 a = {};
    if (tile < ntiles && tile * 16 + col < n) a = X[(size_t)(tile * 16 + col) * 4 + grp];
    for (; tile < ntiles; tile += nwaves) {
        const int nt = tile + nwaves;
        bf16x8 an = {};
        if (nt < ntiles && nt * 16 + col < n) an = X[(size_t)(nt * 16 + col) * 4 + grp];
        const int i = tile * 16 + col;
        const bool writer = grp == 0 && i < n;
        float4 p = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (xa && writer) p = xp[i];
        f32x4 z[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) z[j] = mfma(W1[j], a, zero);
        bf16x8 h[2];
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int e = 0; e < 8; ++e)
                h[s][e] = (__bf16)silu_fast(z[2 * s + e / 4][e % 4] + B1[2 * s + e / 4][e % 4]);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            z[j] = mfma(W2[0][j], h[0], zero);
            z[j] = mfma(W2[1][j], h[1], z[j]);
        }
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int e = 0; e < 8; ++e)
                h[s][e] = (__bf16)silu_fast(z[2 * s + e / 4][e % 4] + B2[2 * s + e / 4][e % 4]);
        f32x4 o = mfma(W3[0], h[0], zero);
        o = mfma(W3[1], h[1], o);
        if (writer) {
            const float dx = (o[0] + B3[0]) * out_scale;
            const float dy = (o[1] + B3[1]) * out_scale;
            const float dz = (o[2] + B3[2]) * out_scale;
            if (out) {
                out[(size_t)i * 3 + 0] = dx;
                out[(size_t)i * 3 + 1] = dy;
                out[(size_t)i * 3 + 2] = dz;
            }
            if (xa) {
                const float* lo = k.lo;
                const float* hi = k.hi;
                if (k.scenes) {
                    const SceneBox& sb = k.scenes[k.sid_out[i]];
                    lo = sb.lo;
                    hi = sb.hi;
                }
                xa[i] = make_float4(rclamp(p.x + dx, lo[0], hi[0]), rclamp(p.y + dy, lo[1], hi[1]),
                                    rclamp(p.z + dz, lo[2], hi[2]), 0.0f);
            }
        }
        a = an;
    }
}

hipError_t launch_mlp_forward(const void* X, int n, const float* params, const void* packed,
                              float out_scale, float* out, const float4* xp, float4* xa,
                              const PbfK& k, hipStream_t s) {
    const int ntiles = (n + 15) / 16;
    int blocks = (ntiles + kMlpWaves - 1) / kMlpWaves;
    if (blocks > 256 * 16) blocks = 256 * 16;
    if (blocks < 1) blocks = 1;
    launch_kernel(k_mlp_forward, dim3(blocks), dim3(64 * kMlpWaves), 0, s, (const bf16x8*)X, n,
                       params, (const bf16x8*)packed, out_scale, out, xp, xa, k);
    return hipGetLastError();
}

// ------------------------------------------------------------------ training
// C[M][N] (fp32, row stride ldc) = sum_k A(m,k) B(k,n) over bf16 operands
// addressed through strides: A(m,k) = A[m*sam + k*sak], B(k,n) = B[k*sbk + n*sbn].
// ones_row: row M-1 of A is all ones (not read), which appends the column
// sums of B -- the bias gradient under the weight gradient.  One wave per
// 16x16 C tile; blockIdx.y takes K in chunks of kchunk (a multiple of 32) and
// writes C + blockIdx.y * M * ldc, summed afterwards in chunk order, so the
// result does not depend on scheduling.
__global__ __launch_bounds__(64) void k_gemm_bf16(const __bf16* __restrict__ A, long sam, long sak,
                                                  const __bf16* __restrict__ B, long sbk, long sbn,
                                                  float* __restrict__ C, int ldc, int M, int N, int K,
                                                  int kchunk, int ones_row) {
    const int lane = threadIdx.x;
    const int col = lane % 16, grp = lane / 16;
    const int tiles_n = (N + 15) / 16;
    const int m0 = (blockIdx.x / tiles_n) * 16, n0 = (blockIdx.x % tiles_n) * 16;
    const int kb = blockIdx.y * kchunk;
    const int ke = min(K, kb + kchunk);
    const int m = m0 + col, nn = n0 + col;
    const bool ones = ones_row && m == M - 1;
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int k0 = kb; k0 < ke; k0 += 32) {
        bf16x8 a, b;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int kk = k0 + 8 * grp + e;
            const bool kin = kk < ke;
            a[e] = (m < M && kin) ? (ones ? (__bf16)1.0f : A[m * sam + kk * sak]) : (__bf16)0.0f;
            b[e] = (nn < N && kin) ? B[kk * sbk + nn * sbn] : (__bf16)0.0f;
        }
        acc = mfma(a, b, acc);
    }
    float* Cc = C + (size_t)blockIdx.y * M * ldc;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int mr = m0 + 4 * grp + i;
        if (mr < M && nn < N) Cc[(size_t)mr * ldc + nn] = acc[i];
    }
}

hipError_t launch_gemm_bf16(const void* A, long sam, long sak, const void* B, long sbk, long sbn,
                            float* C, int ldc, int M, int N, int K, int kchunk, int ones_row,
                            hipStream_t s) {
    if (kchunk <= 0 || kchunk % 32) return hipErrorInvalidValue;
    const int tiles = ((M + 15) / 16) * ((N + 15) / 16);
    const int chunks = (K + kchunk - 1) / kchunk;
    launch_kernel(k_gemm_bf16, dim3(tiles, chunks), dim3(64), 0, s, (const __bf16*)A, sam, sak,
                       (const __bf16*)B, sbk, sbn, C, ldc, M, N, K, kchunk, ones_row);
    return hipGetLastError();
}

// out[t] = sum over chunks c (in order) of part[c * MN + t]
__global__ void k_sum_partials(const float* __restrict__ part, int chunks, int MN, float* __restrict__ out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= MN) return;
    float s = 0.0f;
    for (int c = 0; c < chunks; ++c) s += part[(size_t)c * MN + t];
    out[t] = s;
}

hipError_t launch_sum_partials(const float* part, int chunks, int MN, float* out, hipStream_t s) {
    launch_kernel(k_sum_partials, dim3((MN + 255) / 256), dim3(256), 0, s, part, chunks, MN, out);
    return hipGetLastError();
}

// z[r][c] += bias[c]; h = bf16(silu(z)); z (fp32, with bias) kept for the backward pass
__global__ void k_bias_silu(float* __restrict__ z, const float* __restrict__ bias, int rows, int N,
                            __bf16* __restrict__ h) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)rows * N) return;
    const float v = z[t] + bias[t % N];
    z[t] = v;
    h[t] = (__bf16)silu(v);
}

hipError_t launch_bias_silu(float* z, const float* bias, int rows, int N, void* h, hipStream_t s) {
    const long tot = (long)rows * N;
    launch_kernel(k_bias_silu, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, z, bias, rows, N,
                       (__bf16*)h);
    return hipGetLastError();
}

// Loss L = (1/rows) sum_r sum_c (o_rc - y_rc * inv_scale)^2 with o = z3 + b3.
// dbf = bf16(dL/do); err[r] = sum_c (o_rc - y_rc * inv_scale)^2.
__global__ void k_out_loss(const float* __restrict__ z3, const float* __restrict__ b3,
                           const float* __restrict__ y, int rows, float inv_scale,
                           __bf16* __restrict__ dbf, float* __restrict__ err) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    float e = 0.0f;
    for (int c = 0; c < 3; ++c) {
        const float o = z3[(size_t)r * 3 + c] + b3[c];
        const float diff = o - y[(size_t)r * 3 + c] * inv_scale;
        dbf[(size_t)r * 3 + c] = (__bf16)(2.0f * diff / (float)rows);
        e += diff * diff;
    }
    err[r] = e;
}

hipError_t launch_out_loss(const float* z3, const float* b3, const float* y, int rows, float inv_scale,
                           void* dbf, float* err, hipStream_t s) {
    launch_kernel(k_out_loss, dim3((rows + 255) / 256), dim3(256), 0, s, z3, b3, y, rows, inv_scale,
                       (__bf16*)dbf, err);
    return hipGetLastError();
}

// dz = bf16(dh * silu'(z))
__global__ void k_silu_back(const float* __restrict__ dh, const float* __restrict__ z, long tot,
                            __bf16* __restrict__ dzb) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= tot) return;
    dzb[t] = (__bf16)(dh[t] * silu_grad(z[t]));
}

hipError_t launch_silu_back(const float* dh, const float* z, long tot, void* dzb, hipStream_t s) {
    launch_kernel(k_silu_back, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, dh, z, tot,
                       (__bf16*)dzb);
    return hipGetLastError();
}

__global__ void k_to_bf16(const float* __restrict__ a, long tot, __bf16* __restrict__ b) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < tot) b[t] = (__bf16)a[t];
}

hipError_t launch_to_bf16(const float* a, long tot, void* b, hipStream_t s) {
    launch_kernel(k_to_bf16, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, a, tot, (__bf16*)b);
    return hipGetLastError();
}

// xb[r][c] = bf16((x[src][c] - mu[c]) * inv_sigma[c]), src = rows ? rows[r] : r
__global__ void k_normalize_rows(const float* __restrict__ x, const int* __restrict__ rows, long nrows,
                                 const float* __restrict__ mu, const float* __restrict__ inv_sigma,
                                 __bf16* __restrict__ xb) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nrows * kFeat) return;
    const long r = t / kFeat;
    const int c = (int)(t % kFeat);
    const long src = rows ? rows[r] : r;
    xb[t] = (__bf16)((x[src * kFeat + c] - mu[c]) * inv_sigma[c]);
}

hipError_t launch_normalize_rows(const float* x, const int* rows, long nrows, const float* mu,
                                 const float* inv_sigma, void* xb, hipStream_t s) {
    const long tot = nrows * kFeat;
    launch_kernel(k_normalize_rows, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, x, rows, nrows,
                       mu, inv_sigma, (__bf16*)xb);
    return hipGetLastError();
}

__global__ void k_gather_rows(const float* __restrict__ y, const int* __restrict__ rows, long nrows, int w,
                              float* __restrict__ out) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nrows * w) return;
    const long r = t / w;
    const int c = (int)(t % w);
    out[t] = y[(long)rows[r] * w + c];
}

hipError_t launch_gather_rows(const float* y, const int* rows, long nrows, int w, float* out, hipStream_t s) {
    const long tot = nrows * w;
    launch_kernel(k_gather_rows, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, y, rows, nrows, w,
                       out);
    return hipGetLastError();
}

// *acc += sum of v[0..n) (one block; fixed order per thread, then a tree)
__global__ __launch_bounds__(256) void k_sum_into(const float* __restrict__ v, int n, float* __restrict__ acc) {
    __shared__ float part[256];
    float s = 0.0f;
    for (int i = threadIdx.x; i < n; i += 256) s += v[i];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int d = 128; d >= 1; d >>= 1) {
        if (threadIdx.x < d) part[threadIdx.x] += part[threadIdx.x + d];
        __syncthreads();
    }
    if (threadIdx.x == 0) acc[0] += part[0];
}

hipError_t launch_sum_into(const float* v, int n, float* acc, hipStream_t s) {
    launch_kernel(k_sum_into, dim3(1), dim3(256), 0, s, v, n, acc);
    return hipGetLastError();
}

// out[2c] = sum_r x[r][c], out[2c+1] = sum_r x[r][c]^2 (float64), one block per column
__global__ __launch_bounds__(256) void k_col_stats(const float* __restrict__ x, long rows, int w,
                                                   double* __restrict__ out) {
    __shared__ double s1[256], s2[256];
    const int c = blockIdx.x;
    double a = 0.0, b = 0.0;
    for (long r = threadIdx.x; r < rows; r += 256) {
        const double v = x[r * w + c];
        a += v;
        b += v * v;
    }
    s1[threadIdx.x] = a;
    s2[threadIdx.x] = b;
    __syncthreads();
    for (int d = 128; d >= 1; d >>= 1) {
        if (threadIdx.x < d) {
            s1[threadIdx.x] += s1[threadIdx.x + d];
            s2[threadIdx.x] += s2[threadIdx.x + d];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[2 * c] = s1[0];
        out[2 * c + 1] = s2[0];
    }
}

hipError_t launch_col_stats(const float* x, long rows, int w, double* out, hipStream_t s) {
    launch_kernel(k_col_stats, dim3(w), dim3(256), 0, s, x, rows, w, out);
    return hipGetLastError();
}

// Adam (Kingma & Ba, ICLR 2015) on the fp32 master parameters;
// c1 = 1 / (1 - b1^t), c2 = 1 / (1 - b2^t)
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, int np, float lr, float b1, float b2, float eps, float c1,
                       float c2) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= np) return;
    const float gt = g[t];
    const float mt = b1 * m[t] + (1.0f - b1) * gt;
    const float vt = b2 * v[t] + (1.0f - b2) * gt * gt;
    m[t] = mt;
    v[t] = vt;
    p[t] -= lr * (mt * c1) / (sqrtf(vt * c2) + eps);
}

hipError_t launch_adam(float* p, const float* g, float* m, float* v, int np, float lr, float b1, float b2,
                       float eps, float c1, float c2, hipStream_t s) {
    launch_kernel(k_adam, dim3((np + 255) / 256), dim3(256), 0, s, p, g, m, v, np, lr, b1, b2, eps, c1, c2);
    return hipGetLastError();
}

}  // namespace lus
