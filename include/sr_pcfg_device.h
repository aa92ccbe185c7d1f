/*
 * sr_pcfg_device.h -- C-ABI of the grammar kernels in libsr_hip.so: sampling
 * and the postfix conversions on the GPU, one lane per expression row.
 *
 * Conventions: every array is DEVICE memory, row-major int64 as in
 * sr_pcfg.h (whose host engine gives the same tokens and error codes); the
 * work is enqueued on `stream` (a hipStream_t, NULL = the default stream)
 * and nothing is synchronised.  errors[b] receives the row's code (0 = ok,
 * else the reference's error_codes.h value); a failing row is written as
 * NO_OP (or -1 in the children table).  Rows are at most 128 tokens
 * (HARD_MAX_LENGTH).  Returns 0, -1 for arguments out of range, else the
 * HIP error of the launch.
 */
#ifndef SR_PCFG_DEVICE_H
#define SR_PCFG_DEVICE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* B infix rows (B, max_length) drawn from the CSR grammar, row b on the
 * Philox-4x32-10 stream (key = seed, counter = (block, b)); the leftmost
 * derivation of the reference's sampler (src/pcfg_cpu.cpp:10-119,
 * src/pcfg_cuda.cu:10-119, sample_infix src/pcfg.cpp:572-634) with its
 * restart on overflow.  The host engine draws the same rows with
 * sr_pcfg_sample_philox.  n_symbols = entries of nt_rule_ptr minus one. */
int sr_pcfg_sample_device(const int64_t *rhs_ptr, const int64_t *rhs_concat, const int64_t *nt_rule_ptr,
                          const float *cum_probs, int64_t n_symbols, int64_t start_id, int64_t max_length,
                          int64_t max_tries, int64_t *out, int64_t *errors, uint64_t seed, int64_t B, void *stream);

/* Shunting-yard infix -> postfix with the arity check
 * (src/pcfg.cpp:660-755, src/pcfg_cpu.cpp:261-447, src/pcfg_cuda.cu:299-500);
 * n_prec = entries of the precedence table `prec`. */
int sr_pcfg_infix_to_postfix_device(const int64_t *prec, int64_t n_prec, const int64_t *in, int64_t *out,
                                    int64_t *errors, int64_t lp, int64_t rp, int64_t B, int64_t M, void *stream);

/* Ch (B, M, 2) children table of postfix rows (src/pcfg.cpp:957-986,
 * src/pcfg_cpu.cpp:1066-1143, src/pcfg_cuda.cu:1228-1310). */
int sr_pcfg_get_postfix_children_device(const int64_t *in, int64_t *children, int64_t *errors, int64_t B, int64_t M,
                                        void *stream);

/* Fully parenthesised infix rendering (B, M) -> (B, Mi)
 * (src/pcfg.cpp:790-826, src/pcfg_cpu.cpp:449-635, src/pcfg_cuda.cu:502-714). */
int sr_pcfg_postfix_to_infix_device(const int64_t *in, int64_t *out, int64_t *errors, int64_t lp, int64_t rp,
                                    int64_t B, int64_t M, int64_t Mi, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* SR_PCFG_DEVICE_H */
