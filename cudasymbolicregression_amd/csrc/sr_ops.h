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

// Forward of one non-NO_OP node whose opcode is known to be an operator or a
// constant (leaves VAR / LEARNABLE_CONSTANT / UNKNOWN_LEAF are resolved by the
// caller).  `op` is wave-uniform, so the switch compiles to scalar branches.
template <typename T>
__device__ inline T fwd_node(int op, T a0, T a1) {
    const T nan = Mth<T>::qnan();
    switch (op) {
    case CONST_1: return T(1.0);
    case CONST_2: return T(2.0);
    case CONST_3: return T(3.0);
    case CONST_4: return T(4.0);
    case CONST_5: return T(5.0);
    case PI: return T(3.14159265358979323846);
    case E: return T(2.71828182845904523536);
    case SIN: return Mth<T>::sin_(a0);
    case COS: return Mth<T>::cos_(a0);
    case EXP: return Mth<T>::exp_(a0);
    case LOG: return (a0 <= T(0)) ? nan : Mth<T>::log_(a0);
    case SQUARE: return a0 * a0;
    case SQRT: return (a0 < T(0)) ? nan : Mth<T>::sqrt_(a0);
    case TAN: return Mth<T>::tan_(a0);
    case ARCSIN: return (a0 < T(-1) || a0 > T(1)) ? nan : Mth<T>::asin_(a0);
    case ARCCOS: return (a0 < T(-1) || a0 > T(1)) ? nan : Mth<T>::acos_(a0);
    case ARCTAN: return Mth<T>::atan_(a0);
    case SINH: return Mth<T>::sinh_(a0);
    case COSH: return Mth<T>::cosh_(a0);
    case TANH: return Mth<T>::tanh_(a0);
    case FLOOR: return Mth<T>::floor_(a0);
    case CEIL: return Mth<T>::ceil_(a0);
    case LN: return (a0 <= T(0)) ? nan : Mth<T>::log_(a0);
    case LOG10: return (a0 <= T(0)) ? nan : Mth<T>::log10_(a0);
    case NEG: return -a0;
    case INV: return (a0 == T(0)) ? nan : T(1.0) / a0;
    case CUBE: return a0 * a0 * a0;
    case FOURTH: { T sq = a0 * a0; return sq * sq; }
    case FIFTH: { T sq = a0 * a0; return sq * sq * a0; }
    case ADD: return a0 + a1;
    case SUB: return a0 - a1;
    case MUL: return a0 * a1;
    case DIV: return (a1 == T(0)) ? nan : a0 / a1;
    case POW: return (a0 < T(0)) ? nan : Mth<T>::pow_(a0, a1);
    default: return T(0);
    }
}

// Backward of one operator node (arity >= 1) for one point.
// g is the node's adjoint.  On return *add0 / *add1 tell whether go0 / go1
// are added to the children (the reference skips the add when g == 0, except
// SQRT and POW which always add).  Returns ERR_DIV_ZERO or 0.
template <typename T>
__device__ inline int32_t bwd_node(int op, T g, T a0, T a1, T &go0, bool &add0, T &go1, bool &add1) {
    const T zero = T(0);
    const T nan = Mth<T>::qnan();
    const bool gz = (g == zero);
    add0 = false;
    add1 = false;
    switch (op) {
    case SIN: go0 = g * Mth<T>::cos_(a0); add0 = !gz; return 0;
    case COS: go0 = g * -Mth<T>::sin_(a0); add0 = !gz; return 0;
    case EXP: go0 = g * Mth<T>::exp_(a0); add0 = !gz; return 0;
    case LOG: case LN: go0 = div_w(g, a0); add0 = !gz; return 0;
    case SQUARE: go0 = g * (T(2.0) * a0); add0 = !gz; return 0;
    case SQRT:
        go0 = (a0 < zero) ? nan : (gz ? zero : div_w(g * T(0.5), sqrt_w(a0)));
        add0 = true;
        return 0;
    case TAN: { T c = Mth<T>::cos_(a0); go0 = div_w(g, c * c); add0 = !gz; return 0; }
    case ARCSIN:
        go0 = (a0 < T(-1) || a0 > T(1)) ? nan : div_w(g, sqrt_w(T(1.0) - a0 * a0));
        add0 = !gz;
        return 0;
    case ARCCOS:
        go0 = (a0 < T(-1) || a0 > T(1)) ? nan : -div_w(g, sqrt_w(T(1.0) - a0 * a0));
        add0 = !gz;
        return 0;
    case ARCTAN: go0 = div_w(g, T(1.0) + a0 * a0); add0 = !gz; return 0;
    case SINH: go0 = g * Mth<T>::cosh_(a0); add0 = !gz; return 0;
    case COSH: go0 = g * Mth<T>::sinh_(a0); add0 = !gz; return 0;
    case TANH: { T t = Mth<T>::tanh_(a0); go0 = g * (T(1.0) - t * t); add0 = !gz; return 0; }
    case FLOOR: case CEIL: return 0;
    case LOG10: go0 = div_w(g, a0 * log_w(T(10.0))); add0 = !gz; return 0;
    case NEG: go0 = -g; add0 = !gz; return 0;
    case INV: go0 = (a0 == zero) ? nan : -div_w(g, a0 * a0); add0 = !gz; return 0;
    case CUBE: go0 = g * (T(3.0) * (a0 * a0)); add0 = !gz; return 0;
    case FOURTH: go0 = g * (T(4.0) * (a0 * a0 * a0)); add0 = !gz; return 0;
    case FIFTH: { T sq = a0 * a0; go0 = g * (T(5.0) * (sq * sq)); add0 = !gz; return 0; }
    case ADD: go0 = g; go1 = g; add0 = add1 = !gz; return 0;
    case SUB: go0 = g; go1 = -g; add0 = add1 = !gz; return 0;
    case MUL: go0 = g * a1; go1 = g * a0; add0 = add1 = !gz; return 0;
    case DIV:
        if (gz) return 0;
        add0 = add1 = true;
        if (a1 == zero) { go0 = zero; go1 = zero; return ERR_DIV_ZERO; }
        go0 = div_w(g, a1);
        go1 = -div_w(g * a0, a1 * a1);
        return 0;
    case POW:
        add0 = add1 = true;
        if (a0 < zero) { go0 = nan; go1 = nan; return 0; }
        if (gz) { go0 = zero; go1 = zero; return 0; }
        go0 = g * (pow_w(a0, a1 - T(1.0)) * a1);
        go1 = g * (log_w(a0) * pow_w(a0, a1));
        return 0;
    default: return 0;
    }
}

}  // namespace sr
