// Warm-start MLP kernels (gfx950): BF16 operands on MFMA with FP32
// accumulation, hand-written forward, backward and Adam.
//
// The network maps an expression and its targets to starting constants:
//   input  x = one-hot(Ops) over 64 slots x V opcodes  (+)  g(Y) at 256 points
//   h1 = gelu(x W1^T + b1), h2 = gelu(h1 W2^T + b2), h3 = gelu(h2 W3^T + b3)
//   out = h3 W4^T + b4, masked to the LEARNABLE_CONSTANT slots.
// g(y) = sign(y) * log1p(|y|) (bounded growth for exp-type targets).
//
// Every matrix product is C[M,N] = A[M,K] * Bt[N,K]^T ("NT", both operands
// K-contiguous), which is the natural layout of a linear layer (W stored
// out x in); the backward products use it after explicit transposes.
// Products large enough to fill the GPU with 128x128 blocks use k_gemm_nt:
// one wave computes a 64x64 tile as 2x2 v_mfma_f32_32x32x16_bf16 tiles (each
// fragment feeds two MFMAs) with the next K step's fragments in flight; the
// others use k_gemm_nt_small (one 32x32 tile per wave, a block covering
// 64x64).  Operand fragments are read straight from global memory (16
// contiguous bytes per lane and k-step).
#include "sr_device_info.h"
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>

using sr::num_cus;

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int WAVE = 64;
constexpr int N_OUT = 64;  // output slots of the network (padded expression length)

// Round-to-nearest-even FP32 -> BF16 bits (NaN stays NaN).
__device__ __host__ inline uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7FFFFFFFu) > 0x7F800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ inline float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ inline float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ inline float gelu_grad(float z) {
    const float cdf = 0.5f * (1.0f + erff(z * 0.70710678118654752f));
    const float pdf = 0.39894228040143268f * expf(-0.5f * z * z);
    return cdf + z * pdf;
}

enum Epi : int { EPI_STORE = 0, EPI_BIAS = 1, EPI_BIAS_GELU = 2 };

// C = A * Bt^T (+ bias), M and N multiples of 32, K a multiple of 16; A and
// Bt rows start lda / ldb elements apart (multiples of 8).
// MFMA 32x32x16 lane layout (per wave): lane l holds A[i = l % 32][k0 + 8*(l/32) + e]
// and Bt[j = l % 32][k0 + 8*(l/32) + e] for e = 0..7; accumulator register r of
// lane l is C[8*(r/4) + 4*(l/32) + r%4][l % 32].
// A block of 4 waves computes 128 x 128 outputs, each wave 64 x 64 as 2 x 2
// MFMA tiles, so every fragment loaded feeds two MFMAs; the fragments of the
// next K step are loaded before this step's MFMAs are issued.  Fragment rows
// past M or N (the edges of the last block) are clamped to row M-1 / N-1:
// loaded and multiplied, never stored.
__device__ inline void gemm_store_tile(const f32x16 &acc, int i0, int j0, int lane, int N, int epi,
                                       const float *__restrict__ bias, float *__restrict__ out_f32,
                                       uint16_t *__restrict__ out_bf16, float *__restrict__ pre_f32) {
    const int j = j0 + lane % 32;
    const float bj = (epi != EPI_STORE && bias) ? bias[j] : 0.0f;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int i = i0 + 8 * (q / 4) + 4 * (lane / 32) + q % 4;
        const float z = acc[q] + bj;
        const size_t o = (size_t)i * N + j;
        if (epi == EPI_BIAS_GELU) {
            if (pre_f32) pre_f32[o] = z;
            const float h = gelu(z);
            if (out_f32) out_f32[o] = h;
            if (out_bf16) out_bf16[o] = f2bf(h);
        } else {
            if (out_f32) out_f32[o] = z;
            if (out_bf16) out_bf16[o] = f2bf(z);
        }
    }
}

// The same product with one 32 x 32 tile per wave (a block covers 64 x 64):
// four times the blocks of k_gemm_nt, for products too small to fill the
// GPU with 128 x 128 blocks (the weight gradients, M = N = 512, K = 4096).
__global__ __launch_bounds__(256) void k_gemm_nt_small(const uint16_t *__restrict__ A,
                                                      const uint16_t *__restrict__ Bt, const float *__restrict__ bias,
                                                      int M, int N, int K, int lda, int ldb, int epi,
                                                      float *__restrict__ out_f32, uint16_t *__restrict__ out_bf16,
                                                      float *__restrict__ pre_f32) {
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    const int i0 = blockIdx.y * 64 + (wave / 2) * 32;
    const int j0 = blockIdx.x * 64 + (wave % 2) * 32;
    if (i0 >= M || j0 >= N) return;  // whole wave leaves together
    const int r = lane % 32, kh = (lane / 32) * 8;
    const uint16_t *pa = A + (size_t)(i0 + r) * lda + kh;
    const uint16_t *pb = Bt + (size_t)(j0 + r) * ldb + kh;
    f32x16 acc;
#pragma unroll
    for (int q = 0; q < 16; ++q) acc[q] = 0.0f;
    for (int k0 = 0; k0 < K; k0 += 16) {
        const bf16x8 a = *reinterpret_cast<const bf16x8 *>(pa + k0);
        const bf16x8 b = *reinterpret_cast<const bf16x8 *>(pb + k0);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    }
    gemm_store_tile(acc, i0, j0, lane, N, epi, bias, out_f32, out_bf16, pre_f32);
}

__global__ __launch_bounds__(256) void k_gemm_nt(const uint16_t *__restrict__ A, const uint16_t *__restrict__ Bt,
                                                const float *__restrict__ bias, int M, int N, int K, int lda, int ldb,
                                                int epi,
                                                float *__restrict__ out_f32, uint16_t *__restrict__ out_bf16,
                                                float *__restrict__ pre_f32) {
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    const int i0 = blockIdx.y * 128 + (wave / 2) * 64;
    const int j0 = blockIdx.x * 128 + (wave % 2) * 64;
    if (i0 >= M || j0 >= N) return;  // whole wave leaves together
    const int r = lane % 32, kh = (lane / 32) * 8;
    const int ra = i0 + r, rb = i0 + 32 + r < M ? i0 + 32 + r : M - 1;
    const int ca = j0 + r, cb = j0 + 32 + r < N ? j0 + 32 + r : N - 1;
    const uint16_t *pa0 = A + (size_t)ra * lda + kh, *pa1 = A + (size_t)rb * lda + kh;
    const uint16_t *pb0 = Bt + (size_t)ca * ldb + kh, *pb1 = Bt + (size_t)cb * ldb + kh;
    auto ld = [](const uint16_t *p) { return *reinterpret_cast<const bf16x8 *>(p); };
    f32x16 c00, c01, c10, c11;
#pragma unroll
    for (int q = 0; q < 16; ++q) c00[q] = c01[q] = c10[q] = c11[q] = 0.0f;
    auto step = [&](const bf16x8 &x0, const bf16x8 &x1, const bf16x8 &y0, const bf16x8 &y1) {
        c00 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(x0, y0, c00, 0, 0, 0);
        c01 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(x0, y1, c01, 0, 0, 0);
        c10 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(x1, y0, c10, 0, 0, 0);
        c11 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(x1, y1, c11, 0, 0, 0);
    };
    // two register sets: the loads of one K step are in flight while the
    // other set's MFMAs run
    bf16x8 a0 = ld(pa0), a1 = ld(pa1), b0 = ld(pb0), b1 = ld(pb1);
    int k0 = 0;
    for (; k0 + 32 <= K; k0 += 32) {
        const bf16x8 e0 = ld(pa0 + k0 + 16), e1 = ld(pa1 + k0 + 16), f0 = ld(pb0 + k0 + 16), f1 = ld(pb1 + k0 + 16);
        step(a0, a1, b0, b1);
        const int kn = k0 + 32 < K ? k0 + 32 : k0;  // past the end: reload (in bounds, unused)
        a0 = ld(pa0 + kn);
        a1 = ld(pa1 + kn);
        b0 = ld(pb0 + kn);
        b1 = ld(pb1 + kn);
        step(e0, e1, f0, f1);
    }
    if (k0 < K) step(a0, a1, b0, b1);  // K an odd multiple of 16
    const bool i1 = i0 + 32 < M, j1 = j0 + 32 < N;
    gemm_store_tile(c00, i0, j0, lane, N, epi, bias, out_f32, out_bf16, pre_f32);
    if (j1) gemm_store_tile(c01, i0, j0 + 32, lane, N, epi, bias, out_f32, out_bf16, pre_f32);
    if (i1) gemm_store_tile(c10, i0 + 32, j0, lane, N, epi, bias, out_f32, out_bf16, pre_f32);
    if (i1 && j1) gemm_store_tile(c11, i0 + 32, j0 + 32, lane, N, epi, bias, out_f32, out_bf16, pre_f32);
}

// First layer with a one-hot block in its input: z = zy + b + sum_k WT[k V +
// Ops[b][k]] (rows of the transposed BF16 weights for slots k < slots and
// opcodes in [0, V)), i.e. the product of the one-hot columns with the
// weights without multiplying the zeros; then Z = z and H = gelu(z) (BF16).
// zy holds the product of the other input columns.  One block per row;
// rows b >= B (batch padding) take no one-hot term.
__global__ __launch_bounds__(256) void k_onehot_gelu(const float *__restrict__ zy, const uint16_t *__restrict__ WT,
                                                    const float *__restrict__ bias, const int64_t *__restrict__ ops,
                                                    int M, int64_t B, int slots, int vocab, int H,
                                                    const float *__restrict__ noop_suffix, float *__restrict__ pre,
                                                    uint16_t *__restrict__ h_out) {
    __shared__ int rows[64];
    __shared__ int tail;  // first slot of the trailing NO_OP run (noop_suffix)
    const int64_t b = blockIdx.x;
    if (threadIdx.x < 64) {  // wave 0
        const int k = threadIdx.x;
        int r = -1;
        bool op = false;
        if (b < B && k < slots && k < M) {
            const int64_t c = ops[b * M + k];
            if (c >= 0 && c < vocab) r = k * vocab + (int)c;
            op = c != 0;
        }
        rows[k] = r;
        const uint64_t m = __ballot(op);
        if (k == 0) tail = m ? 64 - __clzll(m) : 0;
    }
    __syncthreads();
    const bool sfx = noop_suffix != nullptr && b < B;
    const int kend = sfx ? tail : 64;
    for (int j = threadIdx.x; j < H; j += blockDim.x) {
        float z = zy[(size_t)b * H + j] + bias[j];
        for (int k = 0; k < kend; ++k) {
            const int r = rows[k];
            if (r >= 0) z += bf2f(WT[(size_t)r * H + j]);
        }
        if (sfx) z += noop_suffix[(size_t)tail * H + j];
        pre[(size_t)b * H + j] = z;
        h_out[(size_t)b * H + j] = f2bf(gelu(z));
    }
}

// S[L][j] = sum over slots k = lim-1 down to L of the NO_OP row WT[k*vocab][j]
// (L = 0 .. lim; S[lim][j] = 0): the one-hot term of a trailing NO_OP run.
__global__ void k_noop_suffix(const uint16_t *__restrict__ WT, int lim, int vocab, int H, float *__restrict__ S) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= H) return;
    float acc = 0.0f;
    S[(size_t)lim * H + j] = 0.0f;
    for (int k = lim - 1; k >= 0; --k) {
        acc += bf2f(WT[(size_t)k * vocab * H + j]);
        S[(size_t)k * H + j] = acc;
    }
}

// out[c][r] = in[r][c] for a rows x cols BF16 matrix (32 x 32 tiles in LDS).
__global__ __launch_bounds__(256) void k_transpose(const uint16_t *__restrict__ in, uint16_t *__restrict__ out,
                                                  int rows, int cols) {
    __shared__ uint16_t tile[32][33];
    const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
    const int tx = threadIdx.x % 32, ty = threadIdx.x / 32;  // 32 x 8
    for (int y = ty; y < 32; y += 8) {
        const int r = r0 + y, c = c0 + tx;
        tile[y][tx] = (r < rows && c < cols) ? in[(size_t)r * cols + c] : (uint16_t)0;
    }
    __syncthreads();
    for (int y = ty; y < 32; y += 8) {
        const int c = c0 + y, r = r0 + tx;
        if (c < cols && r < rows) out[(size_t)c * rows + r] = tile[tx][y];
    }
}

// dZ = dH * gelu'(Z) (FP32 in, BF16 and FP32 out).
__global__ void k_gelu_bwd(const float *__restrict__ dH, const float *__restrict__ Z, int64_t n,
                           float *__restrict__ dZ_f32, uint16_t *__restrict__ dZ_bf16) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float d = dH[i] * gelu_grad(Z[i]);
    if (dZ_f32) dZ_f32[i] = d;
    if (dZ_bf16) dZ_bf16[i] = f2bf(d);
}

// db[n] = sum_m dZ[m][n]: FP64 partial sums over chunks of COL_ROWS rows,
// then the chunks in order (deterministic).
constexpr int COL_ROWS = 64;
__global__ void k_col_sum_part(const float *__restrict__ dZ, int M, int N, double *__restrict__ part) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const int r0 = blockIdx.y * COL_ROWS, r1 = min(r0 + COL_ROWS, M);
    double s = 0.0;
    for (int m = r0; m < r1; ++m) s += (double)dZ[(size_t)m * N + n];
    part[(size_t)blockIdx.y * N + n] = s;
}
__global__ void k_col_sum_fin(const double *__restrict__ part, int chunks, int N, float *__restrict__ db) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    double s = 0.0;
    for (int c = 0; c < chunks; ++c) s += part[(size_t)c * N + n];
    db[n] = (float)s;
}

// Adam on FP32 master weights; also refreshes their BF16 copy.
__global__ void k_adam(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, int64_t n, float lr, float b1, float b2, float eps, float bc1,
                       float bc2, float wd, uint16_t *__restrict__ w_bf16) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i] + wd * w[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
    w[i] = wi;
    if (w_bf16) w_bf16[i] = f2bf(wi);
}

__global__ void k_to_bf16(const float *__restrict__ x, int64_t n, uint16_t *__restrict__ y) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = f2bf(x[i]);
}

// Input rows: one-hot over (slot, opcode) for the first `slots` nodes
// (opcodes outside [0, vocab) encode nothing), then g(Y[b][p]) for p <
// n_pts with g(y) = sign(y) log1p|y| (NaN -> 0); with std_block, then the
// standardised targets (y - mean) / std clipped to [-10, 10] (0 when std is
// below 1e-12) and the two scalars g(mean), log1p(std), where mean and std
// are over the n_pts targets with non-finite values read as 0.  Zero padding
// up to row_len; rows b >= B (batch padding up to a multiple of 32) are all
// zero.  One block of 256 threads per row; the statistics are FP64 sums
// reduced in a fixed order.
__global__ __launch_bounds__(256) void k_features(const int64_t *__restrict__ ops, int M, int64_t B, int slots,
                                                 int vocab, const float *__restrict__ Y, int64_t ldy, int n_pts,
                                                 int std_block, int row_len, uint16_t *__restrict__ out) {
    __shared__ double red[2][256];
    const int64_t b = blockIdx.x;
    uint16_t *row = out + (size_t)b * row_len;
    const int onehot = slots * vocab;
    const float *yr = Y + (b < B ? b : 0) * ldy;
    auto clean = [](float y) { return (y == y && fabsf(y) <= 3.0e38f) ? y : 0.0f; };
    double mean = 0.0, sd = 0.0;
    if (std_block) {
        double s1 = 0.0, s2 = 0.0;
        if (b < B)
            for (int p = threadIdx.x; p < n_pts; p += 256) {
                const double y = clean(yr[p]);
                s1 += y;
                s2 += y * y;
            }
        red[0][threadIdx.x] = s1;
        red[1][threadIdx.x] = s2;
        __syncthreads();
        for (int w = 128; w >= 1; w >>= 1) {
            if ((int)threadIdx.x < w) {
                red[0][threadIdx.x] += red[0][threadIdx.x + w];
                red[1][threadIdx.x] += red[1][threadIdx.x + w];
            }
            __syncthreads();
        }
        const double n = n_pts > 0 ? (double)n_pts : 1.0;
        mean = red[0][0] / n;
        const double var = red[1][0] / n - mean * mean;
        sd = var > 0.0 ? sqrt(var) : 0.0;
    }
    const float inv_sd = sd > 1e-12 ? (float)(1.0 / sd) : 0.0f;
    for (int i = threadIdx.x; i < row_len; i += blockDim.x) {
        float v = 0.0f;
        if (b < B) {
            if (i < onehot) {
                const int k = i / vocab, c = i % vocab;
                v = (k < M && ops[b * M + k] == c) ? 1.0f : 0.0f;
            } else if (i < onehot + n_pts) {
                const float y = yr[i - onehot];
                v = (y == y) ? copysignf(log1pf(fminf(fabsf(y), 3.0e38f)), y) : 0.0f;
            } else if (std_block && i < onehot + 2 * n_pts) {
                const float z = (clean(yr[i - onehot - n_pts]) - (float)mean) * inv_sd;
                v = fminf(fmaxf(z, -10.0f), 10.0f);
            } else if (std_block && i == onehot + 2 * n_pts) {
                const float m = (float)mean;
                v = copysignf(log1pf(fabsf(m)), m);
            } else if (std_block && i == onehot + 2 * n_pts + 1) {
                v = log1pf((float)sd);
            }
        }
        row[i] = f2bf(v);
    }
}

// Masked MSE head: d = (out - label) on the constant slots (Ops == 1) of rows
// b < rows, 0 elsewhere (padding rows up to rows_p included);
// dO = 2 d * inv_count (FP32 and BF16), loss[b] = sum_k d^2.
// Block (64, 4): one wave per row, lane k = output slot k (N_OUT == WAVE).
__global__ void k_mse_head(const float *__restrict__ out, const float *__restrict__ label,
                           const int64_t *__restrict__ ops, int M, int64_t rows, int64_t rows_p, float inv_count,
                           float *__restrict__ dO_f32, uint16_t *__restrict__ dO_bf16, float *__restrict__ loss) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.y + threadIdx.y;
    if (b >= rows_p) return;
    const int k = threadIdx.x;
    const bool is_c = b < rows && k < M && ops[b * M + k] == 1;
    const size_t o = (size_t)b * N_OUT + k;
    const float d = is_c ? out[o] - label[b * M + k] : 0.0f;
    const float g = 2.0f * d * inv_count;
    if (dO_f32) dO_f32[o] = g;
    if (dO_bf16) dO_bf16[o] = f2bf(g);
    if (loss) {
        float s = d * d;
        for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off, 64);
        if (k == 0) loss[b] = s;
    }
}

// Predicted constants: C[b][k] = out[b][k] on the constant slots, 0 elsewhere.
__global__ void k_predict(const float *__restrict__ out, const int64_t *__restrict__ ops, int M, int64_t B,
                          float *__restrict__ C) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B * M) return;
    const int64_t b = i / M;
    const int k = (int)(i % M);
    C[i] = ops[i] == 1 ? out[b * N_OUT + k] : 0.0f;
}

// Start point per problem: the warm start where its loss is finite and not
// above the random start's loss, else the random start.  used[b] = 1 when
// the warm start was taken.
__global__ void k_select_start(const float *__restrict__ C_warm, const float *__restrict__ C_rand,
                               const double *__restrict__ loss_warm, const double *__restrict__ loss_rand, int M,
                               int64_t B, float *__restrict__ C_start, int32_t *__restrict__ used) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B * M) return;
    const int64_t b = i / M;
    const double lw = loss_warm[b], lr = loss_rand[b];
    const bool warm = (lw == lw) && !(lw > lr);  // NaN warm loss never wins; NaN random loss always loses
    C_start[i] = warm ? C_warm[i] : C_rand[i];
    if (used && i % M == 0) used[b] = warm ? 1 : 0;
}

inline int last() { return (int)hipGetLastError(); }
inline unsigned blocks(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

}  // namespace

extern "C" {

int sr_mlp_gemm_nt(const void *A, const void *Bt, const float *bias, int64_t M, int64_t N, int64_t K, int64_t lda,
                   int64_t ldb, int epi, float *out_f32, void *out_bf16, float *pre_f32, void *stream) {
    if (M <= 0 || N <= 0 || K <= 0 || M % 32 || N % 32 || K % 16 || epi < 0 || epi > 2) return -1;
    if (lda < K || ldb < K || lda % 8 || ldb % 8 || M > INT32_MAX / 2 || N > INT32_MAX / 2) return -1;
    // fragments are read as 16-byte vectors
    if (A == nullptr || Bt == nullptr || ((uintptr_t)A | (uintptr_t)Bt) % 16 != 0) return -1;
    if (out_f32 == nullptr && out_bf16 == nullptr) return -1;
    if (epi != EPI_STORE && bias == nullptr) return -1;
    if (blocks(N, 128) * blocks(M, 128) >= 2 * (int64_t)num_cus())
        hipLaunchKernelGGL(k_gemm_nt, dim3(blocks(N, 128), blocks(M, 128)), dim3(256), 0, (hipStream_t)stream,
                           (const uint16_t *)A, (const uint16_t *)Bt, bias, (int)M, (int)N, (int)K, (int)lda,
                           (int)ldb, epi, out_f32, (uint16_t *)out_bf16, pre_f32);
    else
        hipLaunchKernelGGL(k_gemm_nt_small, dim3(blocks(N, 64), blocks(M, 64)), dim3(256), 0, (hipStream_t)stream,
                           (const uint16_t *)A, (const uint16_t *)Bt, bias, (int)M, (int)N, (int)K, (int)lda,
                           (int)ldb, epi, out_f32, (uint16_t *)out_bf16, pre_f32);
    return last();
}

int sr_mlp_onehot_gelu(const float *zy, const void *WT, const float *bias, const int64_t *ops, int64_t M, int64_t B,
                       int64_t B_pad, int slots, int vocab, int64_t H, const float *noop_suffix, float *pre,
                       void *h_out, void *stream) {
    if (M <= 0 || B < 0 || B_pad < B || slots < 0 || slots > 64 || vocab <= 0 || H <= 0) return -1;
    if (B_pad) hipLaunchKernelGGL(k_onehot_gelu, dim3((unsigned)B_pad), dim3(256), 0, (hipStream_t)stream, zy,
                                  (const uint16_t *)WT, bias, ops, (int)M, B, slots, vocab, (int)H, noop_suffix, pre,
                                  (uint16_t *)h_out);
    return last();
}

int sr_mlp_noop_suffix(const void *WT, int slots, int vocab, int64_t M, int64_t H, float *S, void *stream) {
    if (slots < 0 || slots > 64 || vocab <= 0 || M <= 0 || H <= 0 || H > INT32_MAX) return -1;
    const int lim = slots < M ? slots : (int)M;
    hipLaunchKernelGGL(k_noop_suffix, dim3(blocks(H, 256)), dim3(256), 0, (hipStream_t)stream,
                       (const uint16_t *)WT, lim, vocab, (int)H, S);
    return last();
}

int sr_mlp_transpose(const void *in, void *out, int64_t rows, int64_t cols, void *stream) {
    if (rows <= 0 || cols <= 0) return -1;
    hipLaunchKernelGGL(k_transpose, dim3(blocks(cols, 32), blocks(rows, 32)), dim3(256), 0, (hipStream_t)stream,
                       (const uint16_t *)in, (uint16_t *)out, (int)rows, (int)cols);
    return last();
}

int sr_mlp_gelu_bwd(const float *dH, const float *Z, int64_t n, float *dZ_f32, void *dZ_bf16, void *stream) {
    if (n < 0) return -1;
    if (n) hipLaunchKernelGGL(k_gelu_bwd, dim3(blocks(n, 256)), dim3(256), 0, (hipStream_t)stream, dH, Z, n, dZ_f32,
                              (uint16_t *)dZ_bf16);
    return last();
}

int64_t sr_mlp_col_sum_workspace_bytes(int64_t M, int64_t N) {
    return M <= 0 || N <= 0 ? 0 : ((M + COL_ROWS - 1) / COL_ROWS) * N * (int64_t)sizeof(double);
}

int sr_mlp_col_sum(const float *dZ, int64_t M, int64_t N, float *db, void *workspace, int64_t workspace_bytes,
                   void *stream) {
    if (M <= 0 || N <= 0 || M > INT32_MAX / 2 || N > INT32_MAX / 2) return -1;
    if (workspace_bytes < sr_mlp_col_sum_workspace_bytes(M, N)) return -2;
    const int chunks = (int)((M + COL_ROWS - 1) / COL_ROWS);
    hipLaunchKernelGGL(k_col_sum_part, dim3(blocks(N, 256), chunks), dim3(256), 0, (hipStream_t)stream, dZ, (int)M,
                       (int)N, (double *)workspace);
    hipLaunchKernelGGL(k_col_sum_fin, dim3(blocks(N, 256)), dim3(256), 0, (hipStream_t)stream,
                       (const double *)workspace, chunks, (int)N, db);
    return last();
}

int sr_mlp_adam(float *w, const float *g, float *m, float *v, int64_t n, float lr, float b1, float b2, float eps,
                int64_t step, float weight_decay, void *w_bf16, void *stream) {
    if (n < 0 || step < 1) return -1;
    const float bc1 = 1.0f - powf(b1, (float)step), bc2 = 1.0f - powf(b2, (float)step);
    if (n) hipLaunchKernelGGL(k_adam, dim3(blocks(n, 256)), dim3(256), 0, (hipStream_t)stream, w, g, m, v, n, lr, b1,
                              b2, eps, bc1, bc2, weight_decay, (uint16_t *)w_bf16);
    return last();
}

int sr_mlp_to_bf16(const float *x, int64_t n, void *y, void *stream) {
    if (n < 0) return -1;
    if (n) hipLaunchKernelGGL(k_to_bf16, dim3(blocks(n, 256)), dim3(256), 0, (hipStream_t)stream, x, n,
                              (uint16_t *)y);
    return last();
}

int sr_mlp_features(const int64_t *ops, int64_t M, int64_t B, int64_t B_pad, int slots, int vocab, const float *Y,
                    int64_t ldy, int n_pts, int std_block, int row_len, void *out, void *stream) {
    if (B < 0 || B_pad < B || M <= 0 || slots < 0 || vocab <= 0 || n_pts < 0 || n_pts > ldy ||
        (std_block != 0 && std_block != 1) ||
        (int64_t)slots * vocab + (int64_t)n_pts * (1 + std_block) + 2 * std_block > row_len)
        return -1;
    if (B_pad) hipLaunchKernelGGL(k_features, dim3((unsigned)B_pad), dim3(256), 0, (hipStream_t)stream, ops, (int)M,
                                  B, slots, vocab, Y, ldy, n_pts, std_block, row_len, (uint16_t *)out);
    return last();
}

int sr_mlp_mse_head(const float *out, const float *label, const int64_t *ops, int64_t M, int64_t rows,
                    int64_t rows_p, float inv_count, float *dO_f32, void *dO_bf16, float *loss, void *stream) {
    if (M <= 0 || M > N_OUT || rows < 0 || rows_p < rows) return -1;
    if (rows_p) hipLaunchKernelGGL(k_mse_head, dim3(blocks(rows_p, 4)), dim3(64, 4), 0, (hipStream_t)stream, out,
                                   label, ops, (int)M, rows, rows_p, inv_count, dO_f32, (uint16_t *)dO_bf16, loss);
    return last();
}

int sr_mlp_predict(const float *out, const int64_t *ops, int64_t M, int64_t B, float *C, void *stream) {
    if (M <= 0 || M > N_OUT || B < 0) return -1;
    if (B) hipLaunchKernelGGL(k_predict, dim3(blocks(B * M, 256)), dim3(256), 0, (hipStream_t)stream, out, ops,
                              (int)M, B, C);
    return last();
}

int sr_mlp_select_start(const float *C_warm, const float *C_rand, const double *loss_warm, const double *loss_rand,
                        int64_t M, int64_t B, float *C_start, int32_t *used, void *stream) {
    if (M <= 0 || B < 0) return -1;
    if (B) hipLaunchKernelGGL(k_select_start, dim3(blocks(B * M, 256)), dim3(256), 0, (hipStream_t)stream, C_warm,
                              C_rand, loss_warm, loss_rand, (int)M, B, C_start, used);
    return last();
}

}  // extern "C"
