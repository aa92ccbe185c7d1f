/*
 * sr_pcfg.h -- C-ABI of libsr_host.so: the probabilistic context-free
 * grammar engine (sampling and the postfix / prefix / infix conversions)
 * behind ProbabilisticContextFreeGrammar.
 *
 * Conventions: all pointers are caller-owned HOST memory; rows are
 * independent and processed with OpenMP; per-row error codes are written to
 * errors[b] (0 = ok; the reference's error_codes.h values otherwise, e.g.
 * 303 sampling gave up, 404/405/407 parsing, 501/505-509 infix rendering);
 * the Python layer turns them into the reference's exception texts
 * (src/pcfg.cpp:504-570).  Token ids are the grammar's symbol ids
 * (terminals as the Operator enum, then '(' = 50+n_x, ')' = 51+n_x, then
 * non-terminals; src/pcfg.cpp:44-237).  lp / rp are the '(' / ')' ids;
 * `prec` is the operator-precedence table indexed by token id.
 */
#ifndef SR_PCFG_H
#define SR_PCFG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* perm <- indices of the grammar's rules ordered by left-hand-side id with
 * std::sort, as the reference's grammar compiler orders them
 * (src/pcfg.cpp:362-372). */
void sr_pcfg_sort_rules(const int64_t *lhs, int64_t n, int64_t *perm);

/* Samples B infix token rows (B, max_length), one std::mt19937 per row
 * seeded with seeds[b], uniform_real_distribution<float> rule choices over
 * the CSR grammar (rhs_ptr / rhs_concat per rule, nt_rule_ptr per
 * non-terminal, cum_probs per rule).  Replaces sample_infix's CPU path
 * (src/pcfg.cpp:572-634, src/pcfg_cpu.cpp:10-119). */
void sr_pcfg_sample(const int64_t *rule_lhs, const int64_t *rhs_ptr, const int64_t *rhs_concat,
                    const int64_t *nt_rule_ptr, const float *cum_probs, int64_t start_id, int64_t max_length,
                    int64_t max_tries, int64_t *out, int64_t *errors, const int64_t *seeds, int64_t B);

/* The device sampler's rows on the host (no counterpart in the reference):
 * sample_infix's derivation with row b drawn from the Philox-4x32-10 stream
 * (key = seed, counter = (block, b), words (w >> 8) * 2^-24), the rows
 * sr_pcfg_sample_device (sr_pcfg_device.h) writes for the same seed. */
void sr_pcfg_sample_philox(const int64_t *rhs_ptr, const int64_t *rhs_concat, const int64_t *nt_rule_ptr,
                           const float *cum_probs, int64_t start_id, int64_t max_length, int64_t max_tries,
                           int64_t *out, int64_t *errors, uint64_t seed, int64_t B);

/* One Philox-4x32-10 block: out = philox(ctr[4], key[2]) (Salmon et al.,
 * SC'11; no counterpart in the reference). */
void sr_philox4x32_10(const uint32_t *ctr, const uint32_t *key, uint32_t *out);

/* Shunting-yard infix -> postfix / prefix with the arity validation pass
 * (src/pcfg.cpp:660-755, src/pcfg_cpu.cpp:121-447). */
void sr_pcfg_infix_to_postfix(const int64_t *prec, const int64_t *in, int64_t *out, int64_t *errors, int64_t lp,
                              int64_t rp, int64_t B, int64_t M);
void sr_pcfg_infix_to_prefix(const int64_t *prec, const int64_t *in, int64_t *out, int64_t *errors, int64_t lp,
                             int64_t rp, int64_t B, int64_t M);

/* Fully parenthesised infix rendering, (B,M) -> (B,Mi)
 * (src/pcfg.cpp:790-862, src/pcfg_cpu.cpp:449-810). */
void sr_pcfg_postfix_to_infix(const int64_t *in, int64_t *out, int64_t *errors, int64_t lp, int64_t rp, int64_t B,
                              int64_t M, int64_t Mi);
void sr_pcfg_prefix_to_infix(const int64_t *in, int64_t *out, int64_t *errors, int64_t lp, int64_t rp, int64_t B,
                             int64_t M, int64_t Mi);

/* Prefix <-> postfix reordering (src/pcfg.cpp:864-925, src/pcfg_cpu.cpp:812-993). */
void sr_pcfg_prefix_to_postfix(const int64_t *in, int64_t *out, int64_t *errors, int64_t B, int64_t M);
void sr_pcfg_postfix_to_prefix(const int64_t *in, int64_t *out, int64_t *errors, int64_t B, int64_t M);

/* Parent index per prefix node, and the Ch (B,M,2) children table of
 * postfix rows (src/pcfg.cpp:926-986, src/pcfg_cpu.cpp:995-1143). */
void sr_pcfg_get_prefix_parent(const int64_t *in, int64_t *parent, int64_t *errors, int64_t B, int64_t M);
void sr_pcfg_get_postfix_children(const int64_t *in, int64_t *children, int64_t *errors, int64_t B, int64_t M);

#ifdef __cplusplus
}
#endif

#endif /* SR_PCFG_H */
