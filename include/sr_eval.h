/*
 * sr_eval.h -- C-ABI of libsr_hip.so: batched postfix-expression evaluation
 * (forward, backward, fused root-only) and device-resident batched L-BFGS
 * constant fitting on gfx950 (MI355X).
 *
 * Conventions (every entry point):
 *   - all pointers are caller-owned DEVICE memory unless stated otherwise;
 *     nothing here allocates memory the caller sees;
 *   - `stream` is a hipStream_t passed as void* (torch's current stream);
 *     every call only enqueues work on it and returns without waiting;
 *   - return 0 on success, a hipError_t value on a HIP failure, or
 *       -1  invalid arguments (sizes out of the supported range),
 *       -2  workspace smaller than the matching *_workspace_bytes(),
 *       -3  a launch would need more LDS than one workgroup may have;
 *   - expression-level error codes (error_codes.h of the reference:
 *     101 child >= parent, 102 arity mismatch, 201 gradient into a NO_OP
 *     node, 204 division by zero in the backward pass) are written to device
 *     flags/arrays; the caller turns them into the reference's exception
 *     texts (symbolic_evaluation.cpp:34, :160, :300).
 *
 * Layouts: "mb" = the reference backend's node-first layouts Ops (M,B),
 * Ch (M,B,2), C (M,B) [, K], cache (M,N,B) [, K]; the fused and fit entry
 * points take the public batch-first layouts Ops (B,M), Ch (B,M,2), C (B,M).
 * Opcodes are the reference's Operator enum (include/operators.h:10-125).
 * At most 128 nodes per expression (M <= 128).
 */
#ifndef SR_EVAL_H
#define SR_EVAL_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Validation of every (m, b) node: replaces validate_inputs /
 * validate_inputs_kernel (src/symbolic_evaluation.cpp:7-37,
 * src/evaluation_kernels_cuda.cu:13-49).  *d_first receives
 * ((m*B + b) << 10 | code) of the first failing node in row-major (m, b)
 * order, or all ones when every node is valid. */
int sr_validate_mb(const int64_t *ops, const int64_t *ch, int64_t M, int64_t B, unsigned long long *d_first,
                   void *stream);

/* Full-cache forward: cache[k][n][b] = value of node k of expression b at
 * point n.  Replaces SymbolicEvaluation::forward + the per-node forward
 * kernels (src/symbolic_evaluation.cpp:39-92,
 * src/evaluation_kernels_cuda.cu:51-216; CPU spec
 * src/evaluation_kernels_cpu.cpp:55-219). */
int sr_forward_mb_f32(const float *X, int64_t N, int64_t n_x, const int64_t *ops, const int64_t *ch, const float *C,
                      int64_t M, int64_t B, float *cache, void *stream);
int sr_forward_mb_f64(const double *X, int64_t N, int64_t n_x, const int64_t *ops, const int64_t *ch,
                      const double *C, int64_t M, int64_t B, double *cache, void *stream);

/* The same forward with a workspace (same results, same replaced reference
 * interface): the nodes of each group of 64 consecutive expressions go
 * through a row buffer in the workspace, and the cache is written 64
 * expressions at a time (whole 256-byte runs in FP32).  Groups whose rows do
 * not fit are written as without a workspace, so any size is valid; the
 * recommended size holds rows for a quarter of the cache plus the index
 * arrays (K = 1 for the single-constant-set entry points). */
int64_t sr_forward_mb_workspace_bytes(int64_t M, int64_t N, int64_t B, int64_t K, int is_f64);
int sr_forward_mb_ws_f32(const float *X, int64_t N, int64_t n_x, const int64_t *ops, const int64_t *ch,
                         const float *C, int64_t M, int64_t B, float *cache, void *workspace, int64_t ws_bytes,
                         void *stream);
int sr_forward_mb_ws_f64(const double *X, int64_t N, int64_t n_x, const int64_t *ops, const int64_t *ch,
                         const double *C, int64_t M, int64_t B, double *cache, void *workspace, int64_t ws_bytes,
                         void *stream);

/* Full-cache backward: grad_cache (M,N,B) holds the upstream gradient on
 * entry; its contents on exit are unspecified (scratch: for M <= 64 in FP32
 * it is only read, the adjoints live in LDS); writes grad_C (M,B) and grad_X
 * (N,n_x, nullable).  Replaces SymbolicEvaluation::backward + the backward
 * kernels (src/symbolic_evaluation.cpp:94-168,
 * src/evaluation_kernels_cuda.cu:218-632; CPU spec
 * src/evaluation_kernels_cpu.cpp:222-662).  *d_err receives the largest
 * code (0, 201, 204). */
int64_t sr_backward_mb_workspace_bytes(int64_t B, int64_t N, int64_t n_x, int is_f64);
/* 1 when the backward of M-node programs only reads grad_cache (so a caller
 * may pass its upstream gradient without a scratch copy), else 0. */
int sr_backward_mb_reads_only(int64_t M, int is_f64);
int sr_backward_mb_f32(float *grad_cache, const float *cache, const int64_t *ops, const int64_t *ch, int64_t M,
                       int64_t B, int64_t N, int64_t n_x, float *grad_C, float *grad_X, void *workspace,
                       int64_t ws_bytes, int32_t *d_err, void *stream);
int sr_backward_mb_f64(double *grad_cache, const double *cache, const int64_t *ops, const int64_t *ch, int64_t M,
                       int64_t B, int64_t N, int64_t n_x, double *grad_C, double *grad_X, void *workspace,
                       int64_t ws_bytes, int32_t *d_err, void *stream);

/* The same backward with a second, row workspace (same results, same
 * replaced reference interface): the upstream gradient and the cached values
 * of each group of 64 consecutive expressions are first gathered 64
 * expressions at a time into per-expression rows, which the reverse sweep
 * then reads point-contiguously.  Groups whose rows do not fit are read as
 * without it, so any rows_bytes is valid; the recommended size
 * (sr_backward_mb_rows_workspace_bytes, K = 1 here) holds rows for a quarter
 * of the cache.  grad_cache is only read. */
int64_t sr_backward_mb_rows_workspace_bytes(int64_t M, int64_t N, int64_t B, int64_t K, int is_f64);
int sr_backward_mb_rows_f32(float *grad_cache, const float *cache, const int64_t *ops, const int64_t *ch, int64_t M,
                            int64_t B, int64_t N, int64_t n_x, float *grad_C, float *grad_X, void *workspace,
                            int64_t ws_bytes, void *rows_ws, int64_t rows_bytes, int32_t *d_err, void *stream);
int sr_backward_mb_rows_f64(double *grad_cache, const double *cache, const int64_t *ops, const int64_t *ch,
                            int64_t M, int64_t B, int64_t N, int64_t n_x, double *grad_C, double *grad_X,
                            void *workspace, int64_t ws_bytes, void *rows_ws, int64_t rows_bytes, int32_t *d_err,
                            void *stream);
/* sr_backward_mb_rows_* for a cache that is the output of the forward on X
 * (N,n_x) and C (M,B): the gathered groups recompute their forward values
 * from X and C, with the forward's own per-node arithmetic (the same bits),
 * instead of reading them from the cache, so only grad_cache is gathered;
 * groups without room in the rows still read the cache.  Same results as
 * sr_backward_mb_rows_* on that cache.  (The reference's autograd backward,
 * src/symbolic_evaluation.cpp:94-168, receives the cache its forward saved.) */
int sr_backward_mb_rows_recompute_f32(float *grad_cache, const float *cache, const float *X, const float *C,
                                      const int64_t *ops, const int64_t *ch, int64_t M, int64_t B, int64_t N,
                                      int64_t n_x, float *grad_C, float *grad_X, void *workspace, int64_t ws_bytes,
                                      void *rows_ws, int64_t rows_bytes, int32_t *d_err, void *stream);
int sr_backward_mb_rows_recompute_f64(double *grad_cache, const double *cache, const double *X, const double *C,
                                      const int64_t *ops, const int64_t *ch, int64_t M, int64_t B, int64_t N,
                                      int64_t n_x, double *grad_C, double *grad_X, void *workspace, int64_t ws_bytes,
                                      void *rows_ws, int64_t rows_bytes, int32_t *d_err, void *stream);

/* K constant sets per expression: C (M,B,K), cache (M,N,B,K).  Replaces
 * evaluate_multiple_constant_backend and its kernels
 * (src/symbolic_evaluation.cpp:171-305, src/evaluation_kernels_cuda.cu:761-1364). */
int sr_multi_forward_mb_f32(const float *X, int64_t N, int64_t n_x, const int64_t *ops, const int64_t *ch,
                            const float *C, int64_t M, int64_t B, int64_t K, float *cache, void *stream);
int sr_multi_forward_mb_f64(const double *X, int64_t N, int64_t n_x, const int64_t *ops, const int64_t *ch,
                            const double *C, int64_t M, int64_t B, int64_t K, double *cache, void *stream);
/* With a workspace of sr_forward_mb_workspace_bytes(M, N, B, K, is_f64) bytes
 * (or any other size), as sr_forward_mb_ws_f32. */
int sr_multi_forward_mb_ws_f32(const float *X, int64_t N, int64_t n_x, const int64_t *ops, const int64_t *ch,
                               const float *C, int64_t M, int64_t B, int64_t K, float *cache, void *workspace,
                               int64_t ws_bytes, void *stream);
int sr_multi_forward_mb_ws_f64(const double *X, int64_t N, int64_t n_x, const int64_t *ops, const int64_t *ch,
                               const double *C, int64_t M, int64_t B, int64_t K, double *cache, void *workspace,
                               int64_t ws_bytes, void *stream);
int64_t sr_multi_backward_mb_workspace_bytes(int64_t B, int64_t N, int64_t n_x, int64_t K, int is_f64);
int sr_multi_backward_mb_f32(float *grad_cache, const float *cache, const int64_t *ops, const int64_t *ch,
                             int64_t M, int64_t B, int64_t N, int64_t n_x, int64_t K, float *grad_C, float *grad_X,
                             void *workspace, int64_t ws_bytes, int32_t *d_err, void *stream);
int sr_multi_backward_mb_f64(double *grad_cache, const double *cache, const int64_t *ops, const int64_t *ch,
                             int64_t M, int64_t B, int64_t N, int64_t n_x, int64_t K, double *grad_C,
                             double *grad_X, void *workspace, int64_t ws_bytes, int32_t *d_err, void *stream);
/* With a row workspace of sr_backward_mb_rows_workspace_bytes(M, N, B, K,
 * is_f64) bytes (or any other size), as sr_backward_mb_rows_f32. */
int sr_multi_backward_mb_rows_f32(float *grad_cache, const float *cache, const int64_t *ops, const int64_t *ch,
                                  int64_t M, int64_t B, int64_t N, int64_t n_x, int64_t K, float *grad_C,
                                  float *grad_X, void *workspace, int64_t ws_bytes, void *rows_ws, int64_t rows_bytes,
                                  int32_t *d_err, void *stream);
int sr_multi_backward_mb_rows_f64(double *grad_cache, const double *cache, const int64_t *ops, const int64_t *ch,
                                  int64_t M, int64_t B, int64_t N, int64_t n_x, int64_t K, double *grad_C,
                                  double *grad_X, void *workspace, int64_t ws_bytes, void *rows_ws,
                                  int64_t rows_bytes, int32_t *d_err, void *stream);
/* As sr_backward_mb_rows_recompute_*, with C (M,B,K). */
int sr_multi_backward_mb_rows_recompute_f32(float *grad_cache, const float *cache, const float *X, const float *C,
                                            const int64_t *ops, const int64_t *ch, int64_t M, int64_t B, int64_t N,
                                            int64_t n_x, int64_t K, float *grad_C, float *grad_X, void *workspace,
                                            int64_t ws_bytes, void *rows_ws, int64_t rows_bytes, int32_t *d_err,
                                            void *stream);
int sr_multi_backward_mb_rows_recompute_f64(double *grad_cache, const double *cache, const double *X,
                                            const double *C, const int64_t *ops, const int64_t *ch, int64_t M,
                                            int64_t B, int64_t N, int64_t n_x, int64_t K, double *grad_C,
                                            double *grad_X, void *workspace, int64_t ws_bytes, void *rows_ws,
                                            int64_t rows_bytes, int32_t *d_err, void *stream);

/* Fused root-only forward + backward over batch-first layouts.  Computes
 * what optimize's closure needs (symbolic_torch/optimization.py:63-101:
 * root = #non-NO_OP - 1, loss = mean_n (y - Y)^2, loss.sum().backward())
 * without materialising the (M,N,B) cache:
 *   y_out (B,N); loss_out (B,) FP64 [mode 1]; grad_C (B,M) [!fwd_only];
 *   grad_X (N,n_x) [nullable]; err_b (B,) per-expression code;
 *   flags[0] = max validation code, flags[1] = max backward code.
 * mode 0: dy (B,N) is the upstream gradient of y; mode 1: dy is the target
 * Y and the upstream gradient is (2*(y - Y))*(1/N).  Loss and dL/dC are
 * reduced in FP64. */
int64_t sr_fused_workspace_bytes(int64_t N, int64_t n_x, int64_t B, int64_t M, int want_gx, int is_f64);
int sr_fused_root_f32(const float *X, int64_t N, int64_t n_x, const int64_t *ops, const int64_t *ch, const float *C,
                      int64_t M, int64_t B, const float *dy, int mode, int fwd_only, float *y_out, double *loss_out,
                      float *grad_C, float *grad_X, int32_t *err_b, int32_t *flags, void *workspace,
                      int64_t workspace_bytes, void *stream);
int sr_fused_root_f64(const double *X, int64_t N, int64_t n_x, const int64_t *ops, const int64_t *ch,
                      const double *C, int64_t M, int64_t B, const double *dy, int mode, int fwd_only,
                      double *y_out, double *loss_out, double *grad_C, double *grad_X, int32_t *err_b,
                      int32_t *flags, void *workspace, int64_t workspace_bytes, void *stream);

/* Per-call launch choice of the fused evaluator (HOST memory; no
 * counterpart in the reference, whose evaluator has one launch shape).
 * variant[k]: launch variant of tape-row bucket k (k = 0..4 for programs of
 * at most 8, 16, 32, 64 and 256 tape rows), or -1 = the automatic choice
 * that sr_fused_root_* (no options) makes: variant 0.  In FP32 with at most
 * 4 variables and one shared X, variants 0 and 1 of buckets 0 and 1 are the
 * wide evaluator (programs compiled once by a pre-pass into plan records,
 * four consecutive points per lane per 16-byte tape access); the other
 * variants, and every call that cannot use the wide evaluator (FP64, more
 * variables, the ragged form), run one wave per program with its own
 * compilation, where the automatic choice takes the one-point-per-lane
 * variant when N padded to 64-point tiles is at most 80% of N padded to
 * 128-point ones (N <= 64 or 129 <= N <= 192).
 * Variants differ in kernel, waves per block and points per lane; all give
 * the same y and error codes bit for bit, and loss, grad_C and grad_X up to
 * the order of their sums (over points in FP64; over expressions in FP32
 * partial rows, then FP64).  The workspace of sr_fused_workspace_bytes
 * serves every variant. */
typedef struct sr_fused_options {
    int variant[5];
} sr_fused_options;
/* Number of variants of bucket k (FP32: is_f64 = 0), or -1. */
int sr_fused_num_variants(int bucket, int is_f64);
/* Waves per block and points per lane of one variant; 0 or -1. */
int sr_fused_variant_info(int bucket, int variant, int is_f64, int *waves_per_block, int *points_per_lane);
int sr_fused_root_opts_f32(const float *X, int64_t N, int64_t n_x, const int64_t *ops, const int64_t *ch,
                           const float *C, int64_t M, int64_t B, const float *dy, int mode, int fwd_only,
                           float *y_out, double *loss_out, float *grad_C, float *grad_X, int32_t *err_b,
                           int32_t *flags, const sr_fused_options *opts, void *workspace, int64_t workspace_bytes,
                           void *stream);
int sr_fused_root_opts_f64(const double *X, int64_t N, int64_t n_x, const int64_t *ops, const int64_t *ch,
                           const double *C, int64_t M, int64_t B, const double *dy, int mode, int fwd_only,
                           double *y_out, double *loss_out, double *grad_C, double *grad_X, int32_t *err_b,
                           int32_t *flags, const sr_fused_options *opts, void *workspace, int64_t workspace_bytes,
                           void *stream);

/* Allocation-free form of sr_fused_root_f32, for callers that cannot
 * allocate or pass device memory for a workspace (code under graph
 * capture, tight serving loops).  No counterpart in the reference.  The
 * workspace is a fixed scratch array inside the library, one copy per GPU
 * (sr_fused_noalloc_max_bytes() bytes: 128 MiB, enough for about 380k
 * expressions of at most 32 nodes without grad_X), held until the process
 * ends.  Calls that need more (each expression takes a 320-byte plan
 * record, longer programs use global tapes, grad_X needs partial rows)
 * return -2 without launching.  The call allocates
 * nothing, creates no stream or event, and runs every kernel on `stream`.
 * Every call on one GPU uses the same scratch: calls that overlap on a GPU
 * (several streams or host threads at once) race and give wrong results;
 * serialise them.  Otherwise the arguments and results are those of
 * sr_fused_root_f32. */
int64_t sr_fused_noalloc_max_bytes(void);
int sr_fused_root_noalloc_f32(const float *X, int64_t N, int64_t n_x, const int64_t *ops, const int64_t *ch,
                              const float *C, int64_t M, int64_t B, const float *dy, int mode, int fwd_only,
                              float *y_out, double *loss_out, float *grad_C, float *grad_X, int32_t *err_b,
                              int32_t *flags, void *stream);

/* Ragged form of sr_fused_root_*: expression b has its own n_pts[b] points
 * at X + x_off[b] (rows of n_x values) and its dy (target in mode 1) and y
 * values at offset y_off[b]; loss_out[b] is the mean over its own points.
 * N is the largest n_pts[b] (a bound; the workspace does not depend on it
 * without grad_X).  No grad_X.  No counterpart in the reference, whose
 * evaluation shares one X across the batch. */
int sr_fused_root_ragged_f32(const float *X, int64_t N, int64_t n_x, const int64_t *x_off, const int64_t *y_off,
                             const int32_t *n_pts, const int64_t *ops, const int64_t *ch, const float *C, int64_t M,
                             int64_t B, const float *dy, int mode, int fwd_only, float *y_out, double *loss_out,
                             float *grad_C, int32_t *err_b, int32_t *flags, void *workspace, int64_t workspace_bytes,
                             void *stream);
int sr_fused_root_ragged_f64(const double *X, int64_t N, int64_t n_x, const int64_t *x_off, const int64_t *y_off,
                             const int32_t *n_pts, const int64_t *ops, const int64_t *ch, const double *C, int64_t M,
                             int64_t B, const double *dy, int mode, int fwd_only, double *y_out, double *loss_out,
                             double *grad_C, int32_t *err_b, int32_t *flags, void *workspace, int64_t workspace_bytes,
                             void *stream);

/* Per-call options of sr_lbfgs_fit (HOST memory; no counterpart in the
 * reference).  variant[k]: launch variant of tape-row bucket k (waves per
 * problem x points per lane, see sr_fit_variant_info), -1 = the built-in
 * default.  stationary_exit: 1 (the default) ends a problem whose remaining
 * iterations of LBFGS_batch.step (symbolic_torch/lbfgs_batch.py:266-326)
 * would repeat one accepted zero step: when every stored history pair is
 * exactly zero, the two-loop direction is zero, and if the objective at x
 * then reproduces (f, g) bit for bit, each remaining iteration changes
 * nothing but the history ring, so the problem ends with the results of
 * running them (same C, loss and iteration count max_iter; only the closure
 * count differs).  0 runs every iteration. */
#define SR_FIT_BUCKETS 5
typedef struct sr_fit_options {
    int variant[SR_FIT_BUCKETS];
    int stationary_exit;
} sr_fit_options;

/* Device-resident batched L-BFGS over the LEARNABLE_CONSTANT slots of B
 * batch-first expressions.  Replaces optimize + LBFGS_batch.step
 * (symbolic_torch/optimization.py:9-112, symbolic_torch/lbfgs_batch.py:61-326)
 * with the same options: max_iter, history (1..16), tol_grad, line_search
 * (0 strong Wolfe, 1 backtracking), ls_max_iter, eps.
 *   X (N,n_x) shared, or rows X[x_off[b] ...] per problem (ragged, nullable);
 *   Y (B,N) targets, or Y[y_off[b] ...] with n_pts[b] points (ragged);
 *   stop_loss (B,) optional early stop on the FP64 loss (nullable);
 *   outputs C_out (B,M) (non-constant slots keep C_init, or NaN where the
 *   reference's go NaN), loss_out (B,) FP64 final loss, iters_out (B,),
 *   nclos_out (B,) objective evaluations (nullable), err_out (B,), flags (2);
 *   opts: host pointer, NULL = defaults.
 * n_x <= 2048.  The call keeps no state between calls; concurrent calls
 * from several host threads are safe (on one caller stream they are
 * enqueued one after the other). */
int64_t sr_lbfgs_workspace_bytes(int64_t B, int64_t M);
int sr_lbfgs_fit(const float *X, const float *Y, int64_t N, int64_t n_x, const int64_t *ops, const int64_t *ch,
                 const float *C_init, int64_t M, int64_t B, const int64_t *x_off, const int64_t *y_off,
                 const int32_t *n_pts, const double *stop_loss, float *C_out, double *loss_out, int32_t *iters_out,
                 int32_t *nclos_out, int32_t *err_out, int32_t *flags, int max_iter, int history, double tol_grad,
                 int line_search, int ls_max_iter, double eps, const sr_fit_options *opts, void *workspace,
                 int64_t ws_bytes, void *stream);

/* Multi-start form of sr_lbfgs_fit: `starts` (K >= 1) fits of each of the B
 * problems in one launch, from K start points per problem.  The multi-start
 * counterpart of the reference's multiple-constant evaluation
 * (evaluate_multiple_constant_backend, src/symbolic_evaluation.cpp:171-305,
 * C of shape (M, B, K)), carried through to the fit: the K fits of a problem
 * share its program, targets, points and stop threshold and run as
 * neighbouring work items.
 *   C_init (B,K,M); outputs C_out (B,K,M), loss_out, iters_out, nclos_out,
 *   err_out (B,K); X, Y, x_off, y_off, n_pts, stop_loss per problem as in
 *   sr_lbfgs_fit; workspace sr_lbfgs_workspace_bytes(B, M) (problems, not
 *   fits).  B * K <= 2^31 - 1.  Each fit is the one sr_lbfgs_fit computes
 *   for that problem and start (bit for bit). */
int sr_lbfgs_fit_multi(const float *X, const float *Y, int64_t N, int64_t n_x, const int64_t *ops,
                       const int64_t *ch, const float *C_init, int64_t M, int64_t B, int64_t starts,
                       const int64_t *x_off, const int64_t *y_off, const int32_t *n_pts, const double *stop_loss,
                       float *C_out, double *loss_out, int32_t *iters_out, int32_t *nclos_out, int32_t *err_out,
                       int32_t *flags, int max_iter, int history, double tol_grad, int line_search, int ls_max_iter,
                       double eps, const sr_fit_options *opts, void *workspace, int64_t ws_bytes, void *stream);

/* Launch-variant table of the fit kernels (constant; no counterpart in the
 * reference; read by the autotuner).  Problems are bucketed by tape rows;
 * each bucket has variants (waves per problem x points per lane per
 * dispatch).  Host-only calls; return -1 on an out-of-range bucket or
 * variant. */
int sr_fit_num_buckets(void);
int sr_fit_num_variants(int bucket);
int sr_fit_default_variant(int bucket);
int sr_fit_variant_info(int bucket, int variant, int *waves, int *ppt);

/* Measurement probe (no counterpart in the reference): `blocks` x 256
 * threads each run 8 independent chains of `iters` FP32 operations of one
 * kind, 0 FMA, 1 v_exp_f32, 2 v_log_f32, 3 v_sin_f32, 4 v_rcp_f32,
 * 5 v_sqrt_f32, 6 v_rsq_f32, and write one float each to out
 * (blocks * 256 floats).  Timing a launch gives the issue rate of the kind;
 * rate(FMA) / rate(kind) is its cost in FMA slots.
 * sr_probe_valu_ops_per_thread(iters) = operations per thread. */
int sr_probe_valu_ops_per_thread(int iters);
int sr_probe_valu_rate(int kind, int64_t blocks, int iters, float *out, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* SR_EVAL_H */
