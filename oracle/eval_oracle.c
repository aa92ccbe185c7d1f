/*
 * ORACLE — test infrastructure, not product code.
 *
 * A plain-C, single-threaded restatement of the reference's CPU evaluation
 * algorithm (symbolic_torch), used only by tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg as the checker.  The product never links,
 * loads or calls it.
 *
 * Followed files (relative to the reference root):
 *   include/operators.h:66-125    opcode table, is_valid_op, get_arity
 *   include/operators.h:222-434   domain-checked scalar wrappers (NaN policy)
 *   src/evaluation_kernels_cpu.cpp:8-52     validate_inputs_cpu_impl
 *   src/evaluation_kernels_cpu.cpp:55-219   forward step (per node k)
 *   src/evaluation_kernels_cpu.cpp:222-662  backward step (per node k)
 *   src/evaluation_kernels_cpu.cpp:700-870  multiple-constant forward
 *   src/evaluation_kernels_cpu.cpp:873-1309 multiple-constant backward
 *   symbolic_torch/optimization.py:55-58,87-99  root index and MSE closure
 *
 * FP32 instantiations call the C float routines (sinf, expf, ...): the
 * reference's C++ wrappers resolve to the float overloads, and a build of the
 * reference's own CPU unit agrees bit for bit with them (checked in
 * tests/test_oracle.py).
 *
 * Deterministic choices where the reference is racy (OpenMP atomics):
 *   - validation reports the code of the first failing (m, b) in row-major
 *     order (the reference keeps whichever thread wrote last);
 *   - backward accumulates grad_C over n and grad_X over b in ascending
 *     order, and reports max(code) when several backward errors occur.
 *
 * Parity pins (each function is checked by at least one):
 *   validate / forward / backward / multi-K: bit-equality with the
 *     reference's own CPU unit (oracle/_ref) on seeded random programs, plus
 *     closed forms (C=3.14: C+1, sin(3.14)) and the NO_OP / 201 / 204 cases.
 *   fused root + MSE closure: equality with forward+backward of this file
 *     with a root-only upstream gradient, and central finite differences in
 *     FP64.
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdint.h>
#include <string.h>
#include <stdlib.h>

enum {
    NO_OP = 0, LEARNABLE_CONSTANT = 1, CONST_1 = 2, CONST_2 = 3, CONST_3 = 4, CONST_4 = 5, CONST_5 = 6,
    PI_OP = 7, E_OP = 8,
    SIN = 10, COS = 11, EXP = 12, LOG = 13, SQUARE = 14, SQRT = 15, TAN = 16, ARCSIN = 17, ARCCOS = 18,
    ARCTAN = 19, SINH = 20, COSH = 21, TANH = 22, FLOOR = 23, CEIL = 24, LN = 25, LOG10 = 26, NEG = 27,
    INV = 28, CUBE = 29, FOURTH = 30, FIFTH = 31,
    ADD = 40, SUB = 41, MUL = 42, DIV = 43, POW = 44, VAR_START_ID = 50
};

static int arity_of(int op) {
    if (op >= SIN && op <= FIFTH) return 1;
    if (op >= ADD && op <= POW) return 2;
    return 0;
}

/* ------------------------------------------------------------------ */
/* validation (evaluation_kernels_cpu.cpp:8-52)                        */
/* ------------------------------------------------------------------ */
int32_t oracle_validate(const int64_t *ops, const int64_t *ch, int64_t M, int64_t B) {
    for (int64_t m = 0; m < M; ++m) {
        for (int64_t b = 0; b < B; ++b) {
            int op = (int)ops[m * B + b];
            int actual = 0;
            for (int i = 0; i < 2; ++i) {
                int64_t c = ch[(m * B + b) * 2 + i];
                if (c != -1) {
                    actual++;
                    if (c >= m) return 101;
                }
            }
            if (actual != arity_of(op)) return 102;
        }
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* scalar semantics, instantiated for float and double                 */
/* ------------------------------------------------------------------ */
#define DEFINE_SCALAR(T, SFX, F)                                                                   \
    static T fwd_op_##SFX(int op, T a0, T a1, T cval, const T *xrow, int64_t n_x) {                \
        const T nanv = (T)NAN;                                                                     \
        switch (op) {                                                                              \
        case LEARNABLE_CONSTANT: return cval;                                                      \
        case CONST_1: return (T)1.0;                                                               \
        case CONST_2: return (T)2.0;                                                               \
        case CONST_3: return (T)3.0;                                                               \
        case CONST_4: return (T)4.0;                                                               \
        case CONST_5: return (T)5.0;                                                               \
        case PI_OP: return (T)M_PI;                                                                \
        case E_OP: return (T)M_E;                                                                  \
        case SIN: return F(sin)(a0);                                                               \
        case COS: return F(cos)(a0);                                                               \
        case EXP: return F(exp)(a0);                                                               \
        case LOG: return (a0 <= (T)0) ? nanv : F(log)(a0);                                         \
        case SQUARE: return a0 * a0;                                                               \
        case SQRT: return (a0 < (T)0) ? nanv : F(sqrt)(a0);                                        \
        case TAN: return F(tan)(a0);                                                               \
        case ARCSIN: return (a0 < (T)-1 || a0 > (T)1) ? nanv : F(asin)(a0);                        \
        case ARCCOS: return (a0 < (T)-1 || a0 > (T)1) ? nanv : F(acos)(a0);                        \
        case ARCTAN: return F(atan)(a0);                                                           \
        case SINH: return F(sinh)(a0);                                                             \
        case COSH: return F(cosh)(a0);                                                             \
        case TANH: return F(tanh)(a0);                                                             \
        case FLOOR: return F(floor)(a0);                                                           \
        case CEIL: return F(ceil)(a0);                                                             \
        case LN: return (a0 <= (T)0) ? nanv : F(log)(a0);                                          \
        case LOG10: return (a0 <= (T)0) ? nanv : F(log10)(a0);                                     \
        case NEG: return -a0;                                                                      \
        case INV: return (a0 == (T)0) ? nanv : (T)1.0 / a0;                                        \
        case CUBE: return a0 * a0 * a0;                                                            \
        case FOURTH: { T sq = a0 * a0; return sq * sq; }                                           \
        case FIFTH: { T sq = a0 * a0; return sq * sq * a0; }                                       \
        case ADD: return a0 + a1;                                                                  \
        case SUB: return a0 - a1;                                                                  \
        case MUL: return a0 * a1;                                                                  \
        case DIV: return (a1 == (T)0) ? nanv : a0 / a1;                                            \
        case POW: return (a0 < (T)0) ? nanv : F(pow)(a0, a1);                                      \
        default:                                                                                   \
            if (op >= VAR_START_ID && op < VAR_START_ID + n_x) return xrow[op - VAR_START_ID];     \
            return (T)0;                                                                           \
        }                                                                                          \
    }                                                                                              \
    static T div_w_##SFX(T a, T b) { return (b == (T)0) ? (T)NAN : a / b; }                        \
    static T log_w_##SFX(T a) { return (a <= (T)0) ? (T)NAN : F(log)(a); }                        \
    static T pow_w_##SFX(T a, T b) { return (a < (T)0) ? (T)NAN : F(pow)(a, b); }                  \
    static T sqrt_w_##SFX(T a) { return (a < (T)0) ? (T)NAN : F(sqrt)(a); }                        \
    /* Backward of one node for one (n, b).  Returns 0, 201 or 204.  go0/go1 receive the        */ \
    /* contributions to the children; *has0/*has1 say whether they are added (the reference    */ \
    /* "continue"s on g == 0 for most ops, but SQRT and POW always add).  gconst receives the  */ \
    /* grad_C contribution, gvar the grad_X contribution.                                      */ \
    static int bwd_op_##SFX(int op, T g, T a0, T a1, int64_t n_x, T *go0, int *has0, T *go1, int *has1, \
                            int *has_c, int *var_idx) {                                            \
        *has0 = *has1 = *has_c = 0;                                                                \
        *var_idx = -1;                                                                             \
        const T zero = (T)0;                                                                       \
        if (op == NO_OP) return (g == zero) ? 0 : 201;                                             \
        switch (op) {                                                                              \
        case LEARNABLE_CONSTANT: if (g == zero) return 0; *has_c = 1; return 0;                    \
        case SIN: if (g == zero) return 0; *go0 = g * F(cos)(a0); *has0 = 1; return 0;             \
        case COS: if (g == zero) return 0; *go0 = g * -F(sin)(a0); *has0 = 1; return 0;            \
        case EXP: if (g == zero) return 0; *go0 = g * F(exp)(a0); *has0 = 1; return 0;             \
        case LOG: if (g == zero) return 0; *go0 = div_w_##SFX(g, a0); *has0 = 1; return 0;         \
        case SQUARE: if (g == zero) return 0; *go0 = g * ((T)2.0 * a0); *has0 = 1; return 0;       \
        case SQRT:                                                                                 \
            if (a0 < zero) *go0 = (T)NAN;                                                          \
            else if (g == zero) *go0 = zero;                                                       \
            else *go0 = div_w_##SFX(g * (T)0.5, sqrt_w_##SFX(a0));                                  \
            *has0 = 1; return 0;                                                                   \
        case ADD: if (g == zero) return 0; *go0 = g; *go1 = g; *has0 = *has1 = 1; return 0;        \
        case SUB: if (g == zero) return 0; *go0 = g; *go1 = -g; *has0 = *has1 = 1; return 0;       \
        case MUL: if (g == zero) return 0; *go0 = g * a1; *go1 = g * a0; *has0 = *has1 = 1; return 0; \
        case DIV:                                                                                  \
            if (g == zero) return 0;                                                               \
            *has0 = *has1 = 1;                                                                     \
            if (a1 == zero) { *go0 = zero; *go1 = zero; return 204; }                              \
            *go0 = div_w_##SFX(g, a1);                                                             \
            *go1 = -div_w_##SFX(g * a0, a1 * a1);                                                  \
            return 0;                                                                              \
        case POW:                                                                                  \
            *has0 = *has1 = 1;                                                                     \
            if (a0 < zero) { *go0 = (T)NAN; *go1 = (T)NAN; return 0; }                             \
            if (g == zero) { *go0 = zero; *go1 = zero; return 0; }                                 \
            *go0 = g * (pow_w_##SFX(a0, a1 - (T)1.0) * a1);                                        \
            *go1 = g * (log_w_##SFX(a0) * pow_w_##SFX(a0, a1));                                    \
            return 0;                                                                              \
        case TAN: { if (g == zero) return 0; T c = F(cos)(a0); *go0 = div_w_##SFX(g, c * c); *has0 = 1; return 0; } \
        case ARCSIN:                                                                               \
            if (g == zero) return 0;                                                               \
            *has0 = 1;                                                                             \
            if (a0 < (T)-1 || a0 > (T)1) *go0 = (T)NAN;                                            \
            else *go0 = div_w_##SFX(g, sqrt_w_##SFX((T)1.0 - a0 * a0));                            \
            return 0;                                                                              \
        case ARCCOS:                                                                               \
            if (g == zero) return 0;                                                               \
            *has0 = 1;                                                                             \
            if (a0 < (T)-1 || a0 > (T)1) *go0 = (T)NAN;                                            \
            else *go0 = -div_w_##SFX(g, sqrt_w_##SFX((T)1.0 - a0 * a0));                           \
            return 0;                                                                              \
        case ARCTAN: if (g == zero) return 0; *go0 = div_w_##SFX(g, (T)1.0 + a0 * a0); *has0 = 1; return 0; \
        case SINH: if (g == zero) return 0; *go0 = g * F(cosh)(a0); *has0 = 1; return 0;           \
        case COSH: if (g == zero) return 0; *go0 = g * F(sinh)(a0); *has0 = 1; return 0;           \
        case TANH: { if (g == zero) return 0; T t = F(tanh)(a0); *go0 = g * ((T)1.0 - t * t); *has0 = 1; return 0; } \
        case FLOOR: case CEIL: return 0;                                                           \
        case LN: if (g == zero) return 0; *go0 = div_w_##SFX(g, a0); *has0 = 1; return 0;          \
        case LOG10: if (g == zero) return 0; *go0 = div_w_##SFX(g, a0 * log_w_##SFX((T)10.0)); *has0 = 1; return 0; \
        case NEG: if (g == zero) return 0; *go0 = -g; *has0 = 1; return 0;                         \
        case INV:                                                                                  \
            if (g == zero) return 0;                                                               \
            *has0 = 1;                                                                             \
            *go0 = (a0 == zero) ? (T)NAN : -div_w_##SFX(g, a0 * a0);                               \
            return 0;                                                                              \
        case CUBE: if (g == zero) return 0; *go0 = g * ((T)3.0 * (a0 * a0)); *has0 = 1; return 0;  \
        case FOURTH: if (g == zero) return 0; *go0 = g * ((T)4.0 * (a0 * a0 * a0)); *has0 = 1; return 0; \
        case FIFTH: { if (g == zero) return 0; T sq = a0 * a0; *go0 = g * ((T)5.0 * (sq * sq)); *has0 = 1; return 0; } \
        default:                                                                                   \
            if (g == zero) return 0;                                                               \
            if (op >= VAR_START_ID && op < VAR_START_ID + n_x) *var_idx = op - VAR_START_ID;       \
            return 0;                                                                              \
        }                                                                                          \
    }

#define FNF(name) name##f
#define FND(name) name
DEFINE_SCALAR(float, f32, FNF)
DEFINE_SCALAR(double, f64, FND)

/* ------------------------------------------------------------------ */
/* full-cache forward / backward (M-first layouts of the backend)      */
/* ------------------------------------------------------------------ */
#define DEFINE_FULL(T, SFX)                                                                         \
    void oracle_forward_##SFX(const T *X, int64_t N, int64_t n_x, const int64_t *ops, const int64_t *ch,  \
                              const T *C, int64_t M, int64_t B, T *cache) {                        \
        for (int64_t k = 0; k < M; ++k)                                                            \
            for (int64_t n = 0; n < N; ++n)                                                        \
                for (int64_t b = 0; b < B; ++b) {                                                  \
                    int op = (int)ops[k * B + b];                                                  \
                    T *dst = &cache[(k * N + n) * B + b];                                          \
                    if (op == NO_OP) { *dst = (T)NAN; continue; }                                  \
                    int ar = arity_of(op);                                                         \
                    T a0 = 0, a1 = 0;                                                              \
                    if (ar >= 1) a0 = cache[(ch[(k * B + b) * 2] * N + n) * B + b];                \
                    if (ar == 2) a1 = cache[(ch[(k * B + b) * 2 + 1] * N + n) * B + b];            \
                    *dst = fwd_op_##SFX(op, a0, a1, C[k * B + b], &X[n * n_x], n_x);               \
                }                                                                                  \
    }                                                                                              \
    int32_t oracle_backward_##SFX(T *grad_cache, const T *cache, const int64_t *ops, const int64_t *ch,   \
                                  int64_t M, int64_t B, int64_t N, int64_t n_x, T *grad_C, T *grad_X) { \
        int32_t err = 0;                                                                           \
        for (int64_t k = M - 1; k >= 0; --k)                                                       \
            for (int64_t n = 0; n < N; ++n)                                                        \
                for (int64_t b = 0; b < B; ++b) {                                                  \
                    int op = (int)ops[k * B + b];                                                  \
                    T g = grad_cache[(k * N + n) * B + b];                                         \
                    int ar = arity_of(op);                                                         \
                    int64_t c0 = ch[(k * B + b) * 2], c1 = ch[(k * B + b) * 2 + 1];                \
                    T a0 = 0, a1 = 0;                                                              \
                    if (ar >= 1) a0 = cache[(c0 * N + n) * B + b];                                 \
                    if (ar == 2) a1 = cache[(c1 * N + n) * B + b];                                 \
                    T go0 = 0, go1 = 0;                                                            \
                    int h0, h1, hc, vi;                                                            \
                    int e = bwd_op_##SFX(op, g, a0, a1, n_x, &go0, &h0, &go1, &h1, &hc, &vi);       \
                    if (e > err) err = e;                                                          \
                    if (h0) grad_cache[(c0 * N + n) * B + b] += go0;                               \
                    if (h1) grad_cache[(c1 * N + n) * B + b] += go1;                               \
                    if (hc) grad_C[k * B + b] += g;                                                \
                    if (vi >= 0) grad_X[n * n_x + vi] += g;                                        \
                }                                                                                  \
        return err;                                                                                \
    }

DEFINE_FULL(float, f32)
DEFINE_FULL(double, f64)

/* ------------------------------------------------------------------ */
/* multiple constant sets (cache (M,N,B,K), C (M,B,K))                 */
/* ------------------------------------------------------------------ */
#define DEFINE_MULTI(T, SFX)                                                                        \
    void oracle_multi_forward_##SFX(const T *X, int64_t N, int64_t n_x, const int64_t *ops,        \
                                    const int64_t *ch, const T *C, int64_t M, int64_t B, int64_t K, T *cache) { \
        for (int64_t k = 0; k < M; ++k)                                                            \
            for (int64_t n = 0; n < N; ++n)                                                        \
                for (int64_t b = 0; b < B; ++b) {                                                  \
                    int op = (int)ops[k * B + b];                                                  \
                    for (int64_t j = 0; j < K; ++j) {                                              \
                        T *dst = &cache[((k * N + n) * B + b) * K + j];                            \
                        if (op >= VAR_START_ID && op < VAR_START_ID + n_x) { *dst = X[n * n_x + op - VAR_START_ID]; continue; } \
                        if (op == NO_OP) { *dst = (T)NAN; continue; }                              \
                        int ar = arity_of(op);                                                     \
                        T a0 = 0, a1 = 0;                                                          \
                        if (ar >= 1) a0 = cache[((ch[(k * B + b) * 2] * N + n) * B + b) * K + j];  \
                        if (ar == 2) a1 = cache[((ch[(k * B + b) * 2 + 1] * N + n) * B + b) * K + j]; \
                        *dst = fwd_op_##SFX(op, a0, a1, C[(k * B + b) * K + j], &X[n * n_x], n_x); \
                    }                                                                              \
                }                                                                                  \
    }                                                                                              \
    int32_t oracle_multi_backward_##SFX(T *grad_cache, const T *cache, const int64_t *ops,         \
                                        const int64_t *ch, int64_t M, int64_t B, int64_t N, int64_t n_x, \
                                        int64_t K, T *grad_C, T *grad_X) {                         \
        int32_t err = 0;                                                                           \
        for (int64_t k = M - 1; k >= 0; --k)                                                       \
            for (int64_t n = 0; n < N; ++n)                                                        \
                for (int64_t b = 0; b < B; ++b) {                                                  \
                    int op = (int)ops[k * B + b];                                                  \
                    if (op >= VAR_START_ID && op < VAR_START_ID + n_x) {                           \
                        T tot = 0;                                                                 \
                        for (int64_t j = 0; j < K; ++j) tot += grad_cache[((k * N + n) * B + b) * K + j]; \
                        if (tot != (T)0) grad_X[n * n_x + op - VAR_START_ID] += tot;               \
                        continue;                                                                  \
                    }                                                                              \
                    int ar = arity_of(op);                                                         \
                    int64_t c0 = ch[(k * B + b) * 2], c1 = ch[(k * B + b) * 2 + 1];                \
                    for (int64_t j = 0; j < K; ++j) {                                              \
                        T g = grad_cache[((k * N + n) * B + b) * K + j];                           \
                        T a0 = 0, a1 = 0;                                                          \
                        if (ar >= 1) a0 = cache[((c0 * N + n) * B + b) * K + j];                   \
                        if (ar == 2) a1 = cache[((c1 * N + n) * B + b) * K + j];                   \
                        T go0 = 0, go1 = 0;                                                        \
                        int h0, h1, hc, vi;                                                        \
                        int e = bwd_op_##SFX(op, g, a0, a1, n_x, &go0, &h0, &go1, &h1, &hc, &vi);   \
                        if (e > err) err = e;                                                      \
                        if (h0) grad_cache[((c0 * N + n) * B + b) * K + j] += go0;                 \
                        if (h1) grad_cache[((c1 * N + n) * B + b) * K + j] += go1;                 \
                        if (hc) grad_C[(k * B + b) * K + j] += g;                                  \
                    }                                                                              \
                }                                                                                  \
        return err;                                                                                \
    }

DEFINE_MULTI(float, f32)
DEFINE_MULTI(double, f64)

/* ------------------------------------------------------------------ */
/* root-only fused fwd+bwd, batch-first layouts (B,M), (B,M,2), (B,M)  */
/* One expression at a time: forward tape, then a reverse sweep with    */
/* the upstream gradient on the root only.                              */
/*   mode 0: dy (B,N) is the upstream gradient of the root output.      */
/*   mode 1: dy holds the target Y (B,N); upstream = (2*(yhat-Y))*(1/N) */
/*           in the element type (torch's mse_loss/mean backward), and  */
/*           loss_out[b] = mean_n (yhat-Y)^2 accumulated in double.     */
/* root[b] = (#non-NO_OP in row b) - 1, clamped at 0                    */
/* (symbolic_torch/optimization.py:55-58).                              */
/* grad_C is (B,M); grad_X (N,n_x) may be NULL.                         */
/* err_out[b] receives the per-expression code (max over the sweep).    */
/* ------------------------------------------------------------------ */
#define DEFINE_FUSED(T, SFX)                                                                        \
    static void fused_one_##SFX(const T *X, int64_t N, int64_t n_x, const int64_t *ops,            \
                                const int64_t *ch, const T *C, int64_t M, const T *dyb, int mode,  \
                                T *yb, double *loss_b, T *gC, T *grad_X, int32_t *err_b, T *tape,  \
                                T *adj, double *gacc) {                                            \
        const T inv_n = (T)1.0 / (T)N;                                                             \
        int64_t cnt = 0, L = 0;                                                                    \
        for (int64_t k = 0; k < M; ++k)                                                            \
            if (ops[k] != NO_OP) { ++cnt; L = k + 1; }                                             \
        int64_t root = cnt - 1; if (root < 0) root = 0;                                            \
        /* nodes k >= L are NO_OP padding: NaN values nothing reads, zero adjoints */             \
        if (L < root + 1) L = root + 1;                                                            \
        int32_t err = 0;                                                                           \
        double loss = 0.0;                                                                         \
        for (int64_t k = 0; k < M; ++k) gC[k] = 0;                                                 \
        if (gacc) for (int64_t k = 0; k < M; ++k) gacc[k] = 0.0;                                   \
        for (int64_t n = 0; n < N; ++n) {                                                          \
            const T *xrow = &X[n * n_x];                                                           \
            for (int64_t k = 0; k < L; ++k) {                                                      \
                int op = (int)ops[k];                                                              \
                if (op == NO_OP) { tape[k] = (T)NAN; continue; }                                   \
                int ar = arity_of(op);                                                             \
                T a0 = 0, a1 = 0;                                                                  \
                if (ar >= 1) a0 = tape[ch[2 * k]];                                                 \
                if (ar == 2) a1 = tape[ch[2 * k + 1]];                                             \
                tape[k] = fwd_op_##SFX(op, a0, a1, C[k], xrow, n_x);                               \
            }                                                                                      \
            T yh = tape[root];                                                                     \
            yb[n] = yh;                                                                            \
            T g0;                                                                                  \
            if (mode == 1) {                                                                       \
                T d = yh - dyb[n];                                                                 \
                loss += (double)d * (double)d;                                                     \
                g0 = ((T)2.0 * d) * inv_n;                                                         \
            } else g0 = dyb[n];                                                                    \
            for (int64_t k = 0; k < L; ++k) adj[k] = 0;                                            \
            adj[root] = g0;                                                                        \
            for (int64_t k = L - 1; k >= 0; --k) {                                                 \
                int op = (int)ops[k];                                                              \
                int ar = arity_of(op);                                                             \
                int64_t c0 = ch[2 * k], c1 = ch[2 * k + 1];                                        \
                T a0 = 0, a1 = 0;                                                                  \
                if (ar >= 1) a0 = tape[c0];                                                        \
                if (ar == 2) a1 = tape[c1];                                                        \
                T go0 = 0, go1 = 0;                                                                \
                int h0, h1, hc, vi;                                                                \
                int e = bwd_op_##SFX(op, adj[k], a0, a1, n_x, &go0, &h0, &go1, &h1, &hc, &vi);     \
                if (e > err) err = e;                                                              \
                if (h0) adj[c0] += go0;                                                            \
                if (h1) adj[c1] += go1;                                                            \
                if (hc) { if (gacc) gacc[k] += (double)adj[k]; else gC[k] += adj[k]; }             \
                if (vi >= 0 && grad_X) grad_X[n * n_x + vi] += adj[k];                             \
            }                                                                                      \
        }                                                                                          \
        if (gacc) for (int64_t k = 0; k < M; ++k) gC[k] = (T)gacc[k];                              \
        if (loss_b) *loss_b = (mode == 1) ? loss / (double)N : 0.0;                                \
        if (err_b) *err_b = err;                                                                   \
    }                                                                                              \
    /* Expressions are independent: without grad_X (a sum over expressions, */                   \
    /* kept in expression order) they run on all OpenMP threads, with the   */                    \
    /* same results as one after another.  acc64: grad_C summed in double   */                    \
    /* (the device's FP64 reductions) instead of in the element type.       */                    \
    static void fused_root_impl_##SFX(const T *X, int64_t N, int64_t n_x, const int64_t *ops_bm,   \
                                      const int64_t *ch_bm, const T *C_bm, int64_t M, int64_t B,   \
                                      const T *dy, int mode, T *y_out, double *loss_out, T *grad_C, \
                                      T *grad_X, int32_t *err_out, int acc64) {                    \
        _Pragma("omp parallel if (grad_X == NULL && B > 1)")                                       \
        {                                                                                          \
            T *tape = (T *)malloc(sizeof(T) * (size_t)(M > 0 ? M : 1));                            \
            T *adj = (T *)malloc(sizeof(T) * (size_t)(M > 0 ? M : 1));                             \
            double *gacc = acc64 ? (double *)malloc(sizeof(double) * (size_t)(M > 0 ? M : 1)) : NULL; \
            _Pragma("omp for schedule(dynamic, 4)")                                                \
            for (int64_t b = 0; b < B; ++b)                                                        \
                fused_one_##SFX(X, N, n_x, ops_bm + b * M, ch_bm + b * M * 2, C_bm + b * M, M,     \
                                dy ? dy + b * N : NULL, mode, y_out + b * N,                       \
                                loss_out ? loss_out + b : NULL, grad_C + b * M, grad_X,            \
                                err_out ? err_out + b : NULL, tape, adj, gacc);                    \
            free(tape);                                                                            \
            free(adj);                                                                             \
            free(gacc);                                                                            \
        }                                                                                          \
    }                                                                                              \
    void oracle_fused_root_##SFX(const T *X, int64_t N, int64_t n_x, const int64_t *ops_bm,       \
                                 const int64_t *ch_bm, const T *C_bm, int64_t M, int64_t B, const T *dy, \
                                 int mode, T *y_out, double *loss_out, T *grad_C, T *grad_X,        \
                                 int32_t *err_out) {                                               \
        fused_root_impl_##SFX(X, N, n_x, ops_bm, ch_bm, C_bm, M, B, dy, mode, y_out, loss_out,     \
                              grad_C, grad_X, err_out, 0);                                         \
    }                                                                                              \
    void oracle_fused_root_acc64_##SFX(const T *X, int64_t N, int64_t n_x, const int64_t *ops_bm, \
                                       const int64_t *ch_bm, const T *C_bm, int64_t M, int64_t B,  \
                                       const T *dy, int mode, T *y_out, double *loss_out, T *grad_C, \
                                       T *grad_X, int32_t *err_out) {                              \
        fused_root_impl_##SFX(X, N, n_x, ops_bm, ch_bm, C_bm, M, B, dy, mode, y_out, loss_out,     \
                              grad_C, grad_X, err_out, 1);                                         \
    }

DEFINE_FUSED(float, f32)
DEFINE_FUSED(double, f64)
