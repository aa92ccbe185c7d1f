/* sr_mlp.h — C-ABI of the learned warm start in libsr_hip.so (gfx950).
 *
 * No counterpart in the reference: symbolic_torch starts every fit from
 * randn constants (create_constants, symbolic_torch/utils.py:7, used by
 * symbolic_torch/optimization.py:9-112).  These entry points train and run
 * a small MLP whose prediction replaces that start where its loss is lower.
 *
 * Conventions:
 *  - All pointers are device pointers; `stream` is a hipStream_t (NULL =
 *    the default stream).  Calls enqueue work and return without waiting.
 *  - BF16 buffers are passed as `void *` and hold bfloat16 bit patterns
 *    (round to nearest even from FP32, as the kernels write them).
 *  - Return 0 on success, -1 for invalid sizes, -2 for a workspace that is
 *    too small, or a positive hipError_t from the launch.
 */
#ifndef SR_MLP_H
#define SR_MLP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Epilogues of sr_mlp_gemm_nt. */
#define SR_MLP_EPI_STORE 0     /* C = A Bt^T                                 */
#define SR_MLP_EPI_BIAS 1      /* C = A Bt^T + bias[col]                     */
#define SR_MLP_EPI_BIAS_GELU 2 /* Z = A Bt^T + bias, C = gelu(Z) (exact erf) */

/* C[M,N] = A[M,K] * Bt[N,K]^T with BF16 operands (row-major, K contiguous,
 * rows lda / ldb elements apart) and FP32 accumulation on
 * v_mfma_f32_32x32x16_bf16; C is dense (row stride N).  M and N multiples of
 * 32, K a multiple of 16, lda and ldb multiples of 8 and >= K, A and Bt
 * 16-byte aligned (else -1).  Writes whichever of out_f32 (FP32), out_bf16
 * (BF16) are non-NULL (at least one); with EPI_BIAS_GELU, pre_f32 (optional)
 * receives Z.  bias is required by the bias epilogues. */
int sr_mlp_gemm_nt(const void *A, const void *Bt, const float *bias, int64_t M, int64_t N, int64_t K, int64_t lda,
                   int64_t ldb, int epi, float *out_f32, void *out_bf16, float *pre_f32, void *stream);

/* First layer with a one-hot input block, without multiplying its zeros:
 * z[b][j] = zy[b][j] + bias[j] + sum_k WT[k*vocab + Ops[b][k]][j] over
 * slots k < min(slots, M) whose opcode is in [0, vocab) (WT: the BF16
 * weights transposed, one row per input column); pre = z (FP32) and
 * h_out = gelu(z) (BF16), both B_pad x H.  zy (B_pad x H, FP32, may alias
 * pre) is the product of the remaining input columns; rows b >= B take no
 * one-hot term.  With noop_suffix (from sr_mlp_noop_suffix for the same WT,
 * slots, vocab, M and H; nullable) the trailing run of NO_OP slots of a row
 * adds one precomputed sum instead of a gathered row per slot (the padding
 * of short programs); the terms are the same, summed in another order. */
int sr_mlp_onehot_gelu(const float *zy, const void *WT, const float *bias, const int64_t *ops, int64_t M, int64_t B,
                       int64_t B_pad, int slots, int vocab, int64_t H, const float *noop_suffix, float *pre,
                       void *h_out, void *stream);

/* S[L][j] = sum of WT[k*vocab + NO_OP][j] over slots L <= k < min(slots, M),
 * for L = 0 .. min(slots, M) (S: (min(slots, M) + 1) x H FP32). */
int sr_mlp_noop_suffix(const void *WT, int slots, int vocab, int64_t M, int64_t H, float *S, void *stream);

/* out[c][r] = in[r][c] for a rows x cols BF16 matrix. */
int sr_mlp_transpose(const void *in, void *out, int64_t rows, int64_t cols, void *stream);

/* dZ = dH * gelu'(Z), n elements; writes the non-NULL outputs (FP32, BF16). */
int sr_mlp_gelu_bwd(const float *dH, const float *Z, int64_t n, float *dZ_f32, void *dZ_bf16, void *stream);

/* db[n] = sum_m dZ[m][n] over an M x N FP32 matrix, FP64 partial sums in a
 * fixed order.  The workspace holds sr_mlp_col_sum_workspace_bytes(M, N). */
int64_t sr_mlp_col_sum_workspace_bytes(int64_t M, int64_t N);
int sr_mlp_col_sum(const float *dZ, int64_t M, int64_t N, float *db, void *workspace, int64_t workspace_bytes,
                   void *stream);

/* One Adam step on n FP32 parameters (bias-corrected moments, step >= 1,
 * weight_decay added to the gradient); w_bf16 (optional) receives the BF16
 * copy of the updated weights. */
int sr_mlp_adam(float *w, const float *g, float *m, float *v, int64_t n, float lr, float beta1, float beta2,
                float eps, int64_t step, float weight_decay, void *w_bf16, void *stream);

/* y = bf16(x), n elements. */
int sr_mlp_to_bf16(const float *x, int64_t n, void *y, void *stream);

/* Network input rows (BF16, B_pad x row_len): one-hot of Ops[b][k] over
 * opcodes 0..vocab-1 for k < slots, then sign(y) log1p|y| of Y[b][p] for
 * p < n_pts (NaN -> 0); with std_block = 1 also the standardised targets
 * (y - mean) / std clipped to [-10, 10] and the scalars sign(mean)
 * log1p|mean|, log1p(std) (statistics over the n_pts targets, non-finite
 * values read as 0); zeros to row_len; rows B..B_pad-1 all zero.  Ops is
 * (B, M) int64 in the reference layout, Y has row stride ldy. */
int sr_mlp_features(const int64_t *ops, int64_t M, int64_t B, int64_t B_pad, int slots, int vocab, const float *Y,
                    int64_t ldy, int n_pts, int std_block, int row_len, void *out, void *stream);

/* Masked MSE head over network outputs out (rows_p x 64, FP32): for rows
 * b < rows and slots k with Ops[b][k] == LEARNABLE_CONSTANT,
 * d = out[b][k] - label[b][k] (label is (rows, M)), else d = 0.
 * dO = 2 d inv_count (FP32 and/or BF16, rows_p x 64), loss[b] = sum_k d^2. */
int sr_mlp_mse_head(const float *out, const float *label, const int64_t *ops, int64_t M, int64_t rows,
                    int64_t rows_p, float inv_count, float *dO_f32, void *dO_bf16, float *loss, void *stream);

/* C[b][k] = out[b][k] where Ops[b][k] == LEARNABLE_CONSTANT, else 0
 * (C is (B, M) FP32, M <= 64). */
int sr_mlp_predict(const float *out, const int64_t *ops, int64_t M, int64_t B, float *C, void *stream);

/* Per problem b: C_start[b] = C_warm[b] if loss_warm[b] is not NaN and not
 * above loss_rand[b], else C_rand[b]; used[b] (optional) = 1 when the warm
 * start was taken.  Losses are FP64, C arrays (B, M) FP32. */
int sr_mlp_select_start(const float *C_warm, const float *C_rand, const double *loss_warm, const double *loss_rand,
                        int64_t M, int64_t B, float *C_start, int32_t *used, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* SR_MLP_H */
