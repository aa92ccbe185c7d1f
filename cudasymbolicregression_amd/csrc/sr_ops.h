// Operator table and per-node scalar semantics for the gfx950 evaluator.
//
// Opcode values and the domain policy follow the reference's
// include/operators.h:10-125 (enum, get_arity, is_valid_op) and
// include/operators.h:222-434 (NaN policy of every wrapper); the backward
// rules follow src/evaluation_kernels_cpu.cpp:222-662 (the CPU spec: EXP
// recomputes exp(a), CUBE/FOURTH/FIFTH use the CPU rounding order, POW with
// g == 0 and a >= 0 adds +0).  Device math comes from ROCm's device library
// (OCML) through the standard HIP math functions; nothing here uses fast-math
// intrinsics, and the library is compiled with -ffp-contract=off so that
// a*b+c is never fused where the reference rounds twice.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sr {

enum Op : int {
    NO_OP = 0, LEARNABLE_CONSTANT = 1, CONST_1 = 2, CONST_2 = 3, CONST_3 = 4, CONST_4 = 5, CONST_5 = 6,
    PI = 7, E = 8,
    SIN = 10, COS = 11, EXP = 12, LOG = 13, SQUARE = 14, SQRT = 15, TAN = 16, ARCSIN = 17, ARCCOS = 18,
    ARCTAN = 19, SINH = 20, COSH = 21, TANH = 22, FLOOR = 23, CEIL = 24, LN = 25, LOG10 = 26, NEG = 27,
    INV = 28, CUBE = 29, FOURTH = 30, FIFTH = 31,
    ADD = 40, SUB = 41, MUL = 42, DIV = 43, POW = 44, VAR_START_ID = 50,
    // Packed-program code for any opcode that evaluates to 0 as a leaf
    // (unknown ids, variables >= n_x, negative or very large ids).
    UNKNOWN_LEAF = 0xFFFF
};

enum Err : int32_t { ERR_CHILD_GEQ_PARENT = 101, ERR_ARITY = 102, ERR_GRAD_NO_OP = 201, ERR_DIV_ZERO = 204 };

constexpr uint32_t NULL_SLOT = 0xFF;

__host__ __device__ inline int arity_of(int op) {
    if (op >= SIN && op <= FIFTH) return 1;
    if (op >= ADD && op <= POW) return 2;
    return 0;
}

// Packed node word: bits 0..15 opcode (UNKNOWN_LEAF for anything that is a
// 0-valued leaf), 16..23 first child, 24..31 second child (NULL_SLOT = none).
__host__ __device__ inline uint32_t pack_node(int64_t op, int64_t c0, int64_t c1, int64_t n_x) {
    uint32_t code;
    if (op < 0 || op >= UNKNOWN_LEAF) code = UNKNOWN_LEAF;
    else if (op >= VAR_START_ID && op - VAR_START_ID >= n_x) code = UNKNOWN_LEAF;
    else if ((op > E && op < SIN) || (op > FIFTH && op < ADD) || (op > POW && op < VAR_START_ID)) code = UNKNOWN_LEAF;
    else code = (uint32_t)op;
    uint32_t s0 = (c0 < 0 || c0 > 254) ? NULL_SLOT : (uint32_t)c0;
    uint32_t s1 = (c1 < 0 || c1 > 254) ? NULL_SLOT : (uint32_t)c1;
    return code | (s0 << 16) | (s1 << 24);
}
__host__ __device__ inline int node_op(uint32_t w) { return (int)(w & 0xFFFFu); }
__host__ __device__ inline int node_c0(uint32_t w) { return (int)((w >> 16) & 0xFFu); }
__host__ __device__ inline int node_c1(uint32_t w) { return (int)(w >> 24); }

// ---------------------------------------------------------------------------
// Math dispatch per element type (float -> *f routines, double -> plain).
// ---------------------------------------------------------------------------
template <typename T> struct Mth;
template <> struct Mth<float> {
    static __device__ inline float sin_(float a) { return sinf(a); }
    static __device__ inline float cos_(float a) { return cosf(a); }
    static __device__ inline float exp_(float a) { return expf(a); }
    static __device__ inline float log_(float a) { return logf(a); }
    static __device__ inline float sqrt_(float a) { return sqrtf(a); }
    static __device__ inline float tan_(float a) { return tanf(a); }
    static __device__ inline float asin_(float a) { return asinf(a); }
    static __device__ inline float acos_(float a) { return acosf(a); }
    static __device__ inline float atan_(float a) { return atanf(a); }
    static __device__ inline float sinh_(float a) { return sinhf(a); }
    static __device__ inline float cosh_(float a) { return coshf(a); }
    static __device__ inline float tanh_(float a) { return tanhf(a); }
    static __device__ inline float floor_(float a) { return floorf(a); }
    static __device__ inline float ceil_(float a) { return ceilf(a); }
    static __device__ inline float log10_(float a) { return log10f(a); }
    static __device__ inline float pow_(float a, float b) { return powf(a, b); }
    static __device__ inline float qnan() { return __builtin_nanf(""); }
};
template <> struct Mth<double> {
    static __device__ inline double sin_(double a) { return sin(a); }
    static __device__ inline double cos_(double a) { return cos(a); }
    static __device__ inline double exp_(double a) { return exp(a); }
    static __device__ inline double log_(double a) { return log(a); }
    static __device__ inline double sqrt_(double a) { return sqrt(a); }
    static __device__ inline double tan_(double a) { return tan(a); }
    static __device__ inline double asin_(double a) { return asin(a); }
    static __device__ inline double acos_(double a) { return acos(a); }
    static __device__ inline double atan_(double a) { return atan(a); }
    static __device__ inline double sinh_(double a) { return sinh(a); }
    static __device__ inline double cosh_(double a) { return cosh(a); }
    static __device__ inline double tanh_(double a) { return tanh(a); }
    static __device__ inline double floor_(double a) { return floor(a); }
    static __device__ inline double ceil_(double a) { return ceil(a); }
    static __device__ inline double log10_(double a) { return log10(a); }
    static __device__ inline double pow_(double a, double b) { return pow(a, b); }
    static __device__ inline double qnan() { return __builtin_nan(""); }
};

template <typename T> __device__ inline T div_w(T a, T b) { return (b == T(0)) ? Mth<T>::qnan() : a / b; }
template <typename T> __device__ inline T log_w(T a) { return (a <= T(0)) ? Mth<T>::qnan() : Mth<T>::log_(a); }
template <typename T> __device__ inline T sqrt_w(T a) { return (a < T(0)) ? Mth<T>::qnan() : Mth<T>::sqrt_(a); }
template <typename T> __device__ inline T pow_w(T a, T b) { return (a < T(0)) ? Mth<T>::qnan() : Mth<T>::pow_(a, b); }

// Forward of one non-NO_OP node for P points at once (P values per lane),
// whose opcode is an operator or a constant (leaves VAR / LEARNABLE_CONSTANT /
// UNKNOWN_LEAF are resolved by the caller).  `op` is wave-uniform, so the
// switch compiles to scalar branches taken once for all P points.
#define SR_EACH(expr)                                                                              \
    {                                                                                              \
        _Pragma("unroll") for (int p = 0; p < P; ++p) r[p] = (expr);                                \
    }                                                                                              \
    break
template <typename T, int P>
__device__ inline void fwd_nodes(int op, const T *a0, const T *a1, T *r) {
    const T nan = Mth<T>::qnan();
    switch (op) {
    case CONST_1: SR_EACH(T(1.0));
    case CONST_2: SR_EACH(T(2.0));
    case CONST_3: SR_EACH(T(3.0));
    case CONST_4: SR_EACH(T(4.0));
    case CONST_5: SR_EACH(T(5.0));
    case PI: SR_EACH(T(3.14159265358979323846));
    case E: SR_EACH(T(2.71828182845904523536));
    case SIN: SR_EACH(Mth<T>::sin_(a0[p]));
    case COS: SR_EACH(Mth<T>::cos_(a0[p]));
    case EXP: SR_EACH(Mth<T>::exp_(a0[p]));
    case LOG: SR_EACH((a0[p] <= T(0)) ? nan : Mth<T>::log_(a0[p]));
    case SQUARE: SR_EACH(a0[p] * a0[p]);
    case SQRT: SR_EACH((a0[p] < T(0)) ? nan : Mth<T>::sqrt_(a0[p]));
    case TAN: SR_EACH(Mth<T>::tan_(a0[p]));
    case ARCSIN: SR_EACH((a0[p] < T(-1) || a0[p] > T(1)) ? nan : Mth<T>::asin_(a0[p]));
    case ARCCOS: SR_EACH((a0[p] < T(-1) || a0[p] > T(1)) ? nan : Mth<T>::acos_(a0[p]));
    case ARCTAN: SR_EACH(Mth<T>::atan_(a0[p]));
    case SINH: SR_EACH(Mth<T>::sinh_(a0[p]));
    case COSH: SR_EACH(Mth<T>::cosh_(a0[p]));
    case TANH: SR_EACH(Mth<T>::tanh_(a0[p]));
    case FLOOR: SR_EACH(Mth<T>::floor_(a0[p]));
    case CEIL: SR_EACH(Mth<T>::ceil_(a0[p]));
    case LN: SR_EACH((a0[p] <= T(0)) ? nan : Mth<T>::log_(a0[p]));
    case LOG10: SR_EACH((a0[p] <= T(0)) ? nan : Mth<T>::log10_(a0[p]));
    case NEG: SR_EACH(-a0[p]);
    case INV: SR_EACH((a0[p] == T(0)) ? nan : T(1.0) / a0[p]);
    case CUBE: SR_EACH(a0[p] * a0[p] * a0[p]);
    case FOURTH: SR_EACH((a0[p] * a0[p]) * (a0[p] * a0[p]));
    case FIFTH: SR_EACH((a0[p] * a0[p]) * (a0[p] * a0[p]) * a0[p]);
    case ADD: SR_EACH(a0[p] + a1[p]);
    case SUB: SR_EACH(a0[p] - a1[p]);
    case MUL: SR_EACH(a0[p] * a1[p]);
    case DIV: SR_EACH((a1[p] == T(0)) ? nan : a0[p] / a1[p]);
    case POW: SR_EACH((a0[p] < T(0)) ? nan : Mth<T>::pow_(a0[p], a1[p]));
    default: SR_EACH(T(0));
    }
}
#undef SR_EACH

template <typename T>
__device__ inline T fwd_node(int op, T a0, T a1) {
    T r;
    fwd_nodes<T, 1>(op, &a0, &a1, &r);
    return r;
}

// Backward of one operator node (arity >= 1) for P points at once.
// g[p] is the node's adjoint at point p.  On return add0[p] / add1[p] tell
// whether go0[p] / go1[p] are added to the children (the reference skips the
// add when g == 0, except SQRT and POW which always add), and e[p] holds
// ERR_DIV_ZERO or 0.
template <typename T, int P>
__device__ inline void bwd_nodes(int op, const T *g, const T *a0, const T *a1, T *go0, bool *add0, T *go1,
                                 bool *add1, int32_t *e) {
    const T zero = T(0);
    const T nan = Mth<T>::qnan();
#pragma unroll
    for (int p = 0; p < P; ++p) {
        add0[p] = false;
        add1[p] = false;
        e[p] = 0;
        go0[p] = zero;
        go1[p] = zero;
    }
#define SR_EACH(body)                                                                              \
    {                                                                                              \
        _Pragma("unroll") for (int p = 0; p < P; ++p) {                                            \
            const bool gz = (g[p] == zero);                                                        \
            (void)gz;                                                                              \
            body                                                                                   \
        }                                                                                          \
    }                                                                                              \
    break
    switch (op) {
    case SIN: SR_EACH(go0[p] = g[p] * Mth<T>::cos_(a0[p]); add0[p] = !gz;);
    case COS: SR_EACH(go0[p] = g[p] * -Mth<T>::sin_(a0[p]); add0[p] = !gz;);
    case EXP: SR_EACH(go0[p] = g[p] * Mth<T>::exp_(a0[p]); add0[p] = !gz;);
    case LOG:
    case LN: SR_EACH(go0[p] = div_w(g[p], a0[p]); add0[p] = !gz;);
    case SQUARE: SR_EACH(go0[p] = g[p] * (T(2.0) * a0[p]); add0[p] = !gz;);
    case SQRT:
        SR_EACH(go0[p] = (a0[p] < zero) ? nan : (gz ? zero : div_w(g[p] * T(0.5), sqrt_w(a0[p]))); add0[p] = true;);
    case TAN: SR_EACH(const T c = Mth<T>::cos_(a0[p]); go0[p] = div_w(g[p], c * c); add0[p] = !gz;);
    case ARCSIN:
        SR_EACH(go0[p] = (a0[p] < T(-1) || a0[p] > T(1)) ? nan : div_w(g[p], sqrt_w(T(1.0) - a0[p] * a0[p]));
                add0[p] = !gz;);
    case ARCCOS:
        SR_EACH(go0[p] = (a0[p] < T(-1) || a0[p] > T(1)) ? nan : -div_w(g[p], sqrt_w(T(1.0) - a0[p] * a0[p]));
                add0[p] = !gz;);
    case ARCTAN: SR_EACH(go0[p] = div_w(g[p], T(1.0) + a0[p] * a0[p]); add0[p] = !gz;);
    case SINH: SR_EACH(go0[p] = g[p] * Mth<T>::cosh_(a0[p]); add0[p] = !gz;);
    case COSH: SR_EACH(go0[p] = g[p] * Mth<T>::sinh_(a0[p]); add0[p] = !gz;);
    case TANH: SR_EACH(const T t = Mth<T>::tanh_(a0[p]); go0[p] = g[p] * (T(1.0) - t * t); add0[p] = !gz;);
    case FLOOR:
    case CEIL: break;
    case LOG10: SR_EACH(go0[p] = div_w(g[p], a0[p] * log_w(T(10.0))); add0[p] = !gz;);
    case NEG: SR_EACH(go0[p] = -g[p]; add0[p] = !gz;);
    case INV: SR_EACH(go0[p] = (a0[p] == zero) ? nan : -div_w(g[p], a0[p] * a0[p]); add0[p] = !gz;);
    case CUBE: SR_EACH(go0[p] = g[p] * (T(3.0) * (a0[p] * a0[p])); add0[p] = !gz;);
    case FOURTH: SR_EACH(go0[p] = g[p] * (T(4.0) * (a0[p] * a0[p] * a0[p])); add0[p] = !gz;);
    case FIFTH: SR_EACH(const T sq = a0[p] * a0[p]; go0[p] = g[p] * (T(5.0) * (sq * sq)); add0[p] = !gz;);
    case ADD: SR_EACH(go0[p] = g[p]; go1[p] = g[p]; add0[p] = add1[p] = !gz;);
    case SUB: SR_EACH(go0[p] = g[p]; go1[p] = -g[p]; add0[p] = add1[p] = !gz;);
    case MUL: SR_EACH(go0[p] = g[p] * a1[p]; go1[p] = g[p] * a0[p]; add0[p] = add1[p] = !gz;);
    case DIV:
        SR_EACH(if (!gz) {
            add0[p] = add1[p] = true;
            if (a1[p] == zero) {
                e[p] = ERR_DIV_ZERO;
            } else {
                go0[p] = div_w(g[p], a1[p]);
                go1[p] = -div_w(g[p] * a0[p], a1[p] * a1[p]);
            }
        });
    case POW:
        SR_EACH(add0[p] = add1[p] = true; if (a0[p] < zero) {
            go0[p] = nan;
            go1[p] = nan;
        } else if (!gz) {
            go0[p] = g[p] * (pow_w(a0[p], a1[p] - T(1.0)) * a1[p]);
            go1[p] = g[p] * (log_w(a0[p]) * pow_w(a0[p], a1[p]));
        });
    default: break;
    }
#undef SR_EACH
}

// ADD, SUB and MUL (the most frequent operators of typical grammars) are
// evaluated branch-free, ahead of the dispatch tree: one range test instead
// of a chain of opcode comparisons; DIV has its own test next.  Same values
// as fwd_nodes / bwd_nodes.
template <typename T, int P>
__device__ inline void fwd_ops(int op, const T *a0, const T *a1, T *r) {
    if (op >= ADD && op <= MUL) {
#pragma unroll
        for (int p = 0; p < P; ++p) {
            const T s = a0[p] + a1[p], d = a0[p] - a1[p], m = a0[p] * a1[p];
            r[p] = op == MUL ? m : (op == SUB ? d : s);
        }
    } else if (op == DIV) {  // the next most frequent: its own case, not the tree
        fwd_nodes<T, P>(DIV, a0, a1, r);
    } else {
        fwd_nodes<T, P>(op, a0, a1, r);
    }
}
template <typename T, int P>
__device__ inline void bwd_ops(int op, const T *g, const T *a0, const T *a1, T *go0, bool *add0, T *go1, bool *add1,
                               int32_t *e) {
    if (op >= ADD && op <= MUL) {
#pragma unroll
        for (int p = 0; p < P; ++p) {
            const bool mul = op == MUL;
            const T m0 = g[p] * a1[p], m1 = g[p] * a0[p];
            go0[p] = mul ? m0 : g[p];
            go1[p] = mul ? m1 : (op == SUB ? -g[p] : g[p]);
            add0[p] = add1[p] = !(g[p] == T(0));
            e[p] = 0;
        }
    } else if (op == DIV) {
        bwd_nodes<T, P>(DIV, g, a0, a1, go0, add0, go1, add1, e);
    } else {
        bwd_nodes<T, P>(op, g, a0, a1, go0, add0, go1, add1, e);
    }
}

// Opcode of the copy instruction of a leaf root (no reference operator has
// this id).
constexpr int OP_COPY = 63;

// Backward of one instruction for P points with the reference's skip rule
// and the tile group's active points folded in: ge0[p] / ge1[p] are the
// amounts added to the operands' adjoints, 0 where the reference adds
// nothing (or the point is past the end).  Only DIV raises an error (204,
// a zero divisor under a nonzero adjoint at an active point), so only its
// branch touches err; OP_COPY (a leaf root) passes the adjoint through.
// Same values as bwd_nodes with the masks applied.
template <typename T, int P>
__device__ inline void bwd_masked(int op, const T *g, const T *a0, const T *a1, const bool *act, T *ge0, T *ge1,
                                  int32_t &err) {
    if (op >= ADD && op <= MUL) {
#pragma unroll
        for (int p = 0; p < P; ++p) {
            const bool on = act[p] && !(g[p] == T(0));
            const T m0 = g[p] * a1[p], m1 = g[p] * a0[p];
            const T d1 = op == SUB ? -g[p] : g[p];
            ge0[p] = on ? (op == MUL ? m0 : g[p]) : T(0);
            ge1[p] = on ? (op == MUL ? m1 : d1) : T(0);
        }
    } else if (op == OP_COPY) {
#pragma unroll
        for (int p = 0; p < P; ++p) {
            ge0[p] = act[p] ? g[p] : T(0);
            ge1[p] = T(0);
        }
    } else {
        T go0[P], go1[P];
        bool add0[P], add1[P];
        int32_t e[P];
        if (op == DIV) {
            bwd_nodes<T, P>(DIV, g, a0, a1, go0, add0, go1, add1, e);
#pragma unroll
            for (int p = 0; p < P; ++p) err = (act[p] && e[p] > err) ? e[p] : err;
        } else {
            bwd_nodes<T, P>(op, g, a0, a1, go0, add0, go1, add1, e);
        }
#pragma unroll
        for (int p = 0; p < P; ++p) {
            ge0[p] = (act[p] && add0[p]) ? go0[p] : T(0);
            ge1[p] = (act[p] && add1[p]) ? go1[p] : T(0);
        }
    }
}

template <typename T>
__device__ inline int32_t bwd_node(int op, T g, T a0, T a1, T &go0, bool &add0, T &go1, bool &add1) {
    int32_t e;
    bwd_nodes<T, 1>(op, &g, &a0, &a1, &go0, &add0, &go1, &add1, &e);
    return e;
}

}  // namespace sr
