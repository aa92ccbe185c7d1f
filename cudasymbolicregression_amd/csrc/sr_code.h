// Compiled expression programs shared by the fit and the fused evaluator.
//
// A problem's program (node-first words from load_program) is compiled once
// per wave into an instruction list: one instruction per operator node up
// to the root, leaves folded into operand descriptors, plus a copy
// instruction when the root itself is a leaf.  LEARNABLE_CONSTANT nodes are
// COMPACTED: constant j of a program is slot j of an SVec (lane j, and
// 64 + j when SL == 2).  Instruction i writes tape row i, so the root is the
// last row.  Evaluating an instruction is one opcode dispatch for all PPT
// points of a tile group; operand fetch and adjoint scatter are selects on
// the wave-uniform operand descriptors.
#pragma once
#include "sr_wave.h"

namespace sr {

// ------------------------------------------------------------------ slots
template <typename T, int SL>
struct SVec {
    T v[SL];
};

template <typename T, int SL>
__device__ inline T slot_get(const SVec<T, SL> &a, int j) {
    return (SL == 1 || j < WAVE) ? rl(a.v[0], j & (WAVE - 1)) : rl(a.v[SL - 1], j - WAVE);
}

// ---------------------------------------------------------------- program
// One problem's program, node-lane words; LEARNABLE_CONSTANT nodes carry
// their compact ordinal in the first-child field.
template <typename T>
struct CProg {
    uint32_t w0, w1;
    T cinit0, cinit1;       // C of node lane k (the start values for the non-constant slots)
    uint64_t cm0, cm1;      // LEARNABLE_CONSTANT nodes
    int L, root, nc, K;     // K = nc + (M > nc): live slots incl. the shared one
    int32_t verr;
    bool shared_nan_init;   // every non-constant slot starts as NaN
    __device__ inline uint32_t word(int k) const { return k < WAVE ? rl_u32(w0, k) : rl_u32(w1, k - WAVE); }
};

template <typename T>
__device__ inline CProg<T> load_cprog(const int64_t *ops, const int64_t *ch, const T *C, int64_t b, int M, int n_x) {
    Program<T> p = load_program<T>(ops, ch, C, b, M, 1, M, 1, M, n_x);
    CProg<T> f;
    const int lane = lane_id();
    f.cm0 = p.cmask0;
    f.cm1 = p.cmask1;
    f.L = p.L;
    f.root = p.root;
    f.verr = p.verr;
    f.cinit0 = p.c0;
    f.cinit1 = p.c1;
    const int n0 = __popcll(p.cmask0);
    f.nc = n0 + __popcll(p.cmask1);
    f.K = f.nc + (M > f.nc ? 1 : 0);
    const uint64_t below = lane == 0 ? 0ull : (~0ull >> (WAVE - lane));
    const int o0 = __popcll(p.cmask0 & below);
    const int o1 = n0 + __popcll(p.cmask1 & below);
    const uint32_t cw = (uint32_t)LEARNABLE_CONSTANT | (NULL_SLOT << 24);
    f.w0 = ((p.cmask0 >> lane) & 1ull) ? (cw | ((uint32_t)o0 << 16)) : p.w0;
    f.w1 = ((p.cmask1 >> lane) & 1ull) ? (cw | ((uint32_t)o1 << 16)) : p.w1;
    const bool nonconst_finite0 = lane < M && !((p.cmask0 >> lane) & 1ull) && !(p.c0 != p.c0);
    const bool nonconst_finite1 = lane + WAVE < M && !((p.cmask1 >> lane) & 1ull) && !(p.c1 != p.c1);
    f.shared_nan_init = (__ballot(nonconst_finite0) | __ballot(nonconst_finite1)) == 0ull;
    return f;
}

// Compact start point: lane j <- C_init of the j-th constant; shared slot 0.
template <typename T, int SL>
__device__ inline SVec<T, SL> compact_start(const CProg<T> &p) {
    const int lane = lane_id();
    SVec<T, SL> x;
#pragma unroll
    for (int h = 0; h < SL; ++h) x.v[h] = T(0);
    int j = 0;
    for (uint64_t m = p.cm0; m; m &= m - 1, ++j) {
        const int bit = __ffsll((long long)m) - 1;
        const T v = rl(p.cinit0, bit);
        if (j < WAVE) { if (lane == j) x.v[0] = v; }
        else if (lane == j - WAVE) x.v[SL - 1] = v;
    }
    for (uint64_t m = p.cm1; m; m &= m - 1, ++j) {
        const int bit = __ffsll((long long)m) - 1;
        const T v = rl(p.cinit1, bit);
        if (j < WAVE) { if (lane == j) x.v[0] = v; }
        else if (lane == j - WAVE) x.v[SL - 1] = v;
    }
    return x;
}

// ------------------------------------------------------------------ tapes
// ------------------------------------------------------------ instructions
// A problem's program is compiled once into an instruction list: one
// instruction per operator node up to the root (leaves are folded into the
// operand descriptors of their parents), plus a copy instruction when the
// root itself is a leaf.  Instruction i writes tape slot i, so the root is
// always the last slot.  Word: bits 0-5 opcode (the reference's operator id,
// or OP_COPY), bits 6-18 and 19-31 the two operand descriptors
// (2-bit kind, 11-bit index).
constexpr int K_TAPE = 0, K_VAR = 1, K_CONST = 2, K_IMM = 3;
constexpr int IMM_NOOP = 8;  // a NO_OP node used as an operand: NaN, and 201 if it receives a gradient
constexpr int OP_COPY = 63;
constexpr int MAX_NX_CODE = 2047;

__host__ __device__ inline int instr_op(uint32_t w) { return (int)(w & 63u); }
__host__ __device__ inline uint32_t instr_d0(uint32_t w) { return (w >> 6) & 0x1FFFu; }
__host__ __device__ inline uint32_t instr_d1(uint32_t w) { return w >> 19; }
__host__ __device__ inline int desc_kind(uint32_t d) { return (int)(d & 3u); }
__host__ __device__ inline int desc_idx(uint32_t d) { return (int)(d >> 2); }

// Descriptor of node word w as an operand of its parent.
__device__ inline uint32_t operand_desc(uint32_t w, uint32_t instr_index, bool is_instr) {
    const int op = node_op(w);
    if (is_instr) return (uint32_t)K_TAPE | (instr_index << 2);
    if (op == LEARNABLE_CONSTANT) return (uint32_t)K_CONST | ((uint32_t)node_c0(w) << 2);
    if (op >= VAR_START_ID && op != UNKNOWN_LEAF) return (uint32_t)K_VAR | ((uint32_t)(op - VAR_START_ID) << 2);
    if (op >= CONST_1 && op <= E) return (uint32_t)K_IMM | ((uint32_t)(op - 1) << 2);  // 1..5, 6 = pi, 7 = e
    if (op == NO_OP) return (uint32_t)K_IMM | ((uint32_t)IMM_NOOP << 2);
    return (uint32_t)K_IMM;  // unknown leaf: 0
}

// Instruction count of a batch-first program row under the same rule as
// compile_code (used to bucket problems before they are compiled): one wave
// per row, lane k reading node k (and 64 + k), so the reads are coalesced.
// Returns wave-uniform values.
__device__ inline int count_instructions_wave(const int64_t *row, int M, int n_x, int &nc) {
    const int lane = lane_id();
    int64_t op0 = NO_OP, op1 = NO_OP;
    if (lane < M) op0 = row[lane];
    if (lane + WAVE < M) op1 = row[lane + WAVE];
    const uint64_t nn0 = __ballot(op0 != NO_OP), nn1 = __ballot(op1 != NO_OP);
    nc = __popcll(__ballot(op0 == LEARNABLE_CONSTANT)) + __popcll(__ballot(op1 == LEARNABLE_CONSTANT));
    const int nn = __popcll(nn0) + __popcll(nn1);
    const int root = nn > 0 ? nn - 1 : 0;
    const bool is0 = lane < M && lane <= root && arity_of(node_op(pack_node(op0, -1, -1, n_x))) > 0;
    const bool is1 = lane + WAVE < M && lane + WAVE <= root && arity_of(node_op(pack_node(op1, -1, -1, n_x))) > 0;
    const uint64_t m0 = __ballot(is0), m1 = __ballot(is1);
    const int ni = __popcll(m0) + __popcll(m1);
    const bool root_is_op = root < WAVE ? ((m0 >> root) & 1ull) : ((m1 >> (root - WAVE)) & 1ull);
    return ni + (root_is_op ? 0 : 1);
}

struct Code {
    uint32_t iw0, iw1;  // lane i: instruction i (and 64 + i)
    int NI;
    __device__ inline uint32_t word(int i) const { return i < WAVE ? rl_u32(iw0, i) : rl_u32(iw1, i - WAVE); }
    // Instruction word for tapes of at most LT instructions (no second word
    // to choose from when LT <= 64).
    template <int LT>
    __device__ inline uint32_t word_lt(int i) const {
        if constexpr (LT <= WAVE) return rl_u32(iw0, i);
        else return word(i);
    }
};

__device__ inline void place_word(Code &c, int i, uint32_t v) {
    const int lane = lane_id();
    if (i < WAVE) { if (lane == i) c.iw0 = v; }
    else if (lane == i - WAVE) c.iw1 = v;
}

template <typename T>
__device__ inline Code compile_code(const CProg<T> &p, int M) {
    const int lane = lane_id();
    const int root = p.root;
    const uint32_t w0 = p.w0, w1 = p.w1;
    const bool op0 = lane < M && lane <= root && arity_of(node_op(w0)) > 0;
    const bool op1 = lane + WAVE < M && lane + WAVE <= root && arity_of(node_op(w1)) > 0;
    const uint64_t om0 = __ballot(op0), om1 = __ballot(op1);
    const uint64_t below = lane == 0 ? 0ull : (~0ull >> (WAVE - lane));
    const uint32_t ix0 = (uint32_t)__popcll(om0 & below);
    const uint32_t ix1 = (uint32_t)(__popcll(om0) + __popcll(om1 & below));
    const uint32_t d0 = operand_desc(w0, ix0, op0), d1 = operand_desc(w1, ix1, op1);
    auto child_desc = [&](int c) -> uint32_t {  // every lane takes part in both shuffles
        const int cc = c & 127;
        const uint32_t lo = (uint32_t)__shfl((int)d0, cc & (WAVE - 1), WAVE);
        const uint32_t hi = (uint32_t)__shfl((int)d1, cc & (WAVE - 1), WAVE);
        return cc < WAVE ? lo : hi;
    };
    const uint32_t c00 = child_desc(node_c0(w0)), c01 = child_desc(node_c1(w0));
    const uint32_t c10 = child_desc(node_c0(w1)), c11 = child_desc(node_c1(w1));
    const uint32_t i0 = (uint32_t)node_op(w0) | (c00 << 6) | ((arity_of(node_op(w0)) == 2 ? c01 : 0u) << 19);
    const uint32_t i1 = (uint32_t)node_op(w1) | (c10 << 6) | ((arity_of(node_op(w1)) == 2 ? c11 : 0u) << 19);
    Code c;
    c.iw0 = 0u;
    c.iw1 = 0u;
    int i = 0;
    for (uint64_t m = om0; m; m &= m - 1, ++i) place_word(c, i, rl_u32(i0, __ffsll((long long)m) - 1));
    for (uint64_t m = om1; m; m &= m - 1, ++i) place_word(c, i, rl_u32(i1, __ffsll((long long)m) - 1));
    const uint32_t rw = root < WAVE ? rl_u32(w0, root) : rl_u32(w1, root - WAVE);
    if (arity_of(node_op(rw)) == 0) {
        const uint32_t rd = root < WAVE ? rl_u32(d0, root) : rl_u32(d1, root - WAVE);
        place_word(c, i, (uint32_t)OP_COPY | (rd << 6));
        ++i;
    }
    c.NI = i;
    return c;
}

// Tape of one wave: the value of instruction i at point p of a tile group is
// v[(i*PPT + p)*64 + lane], its adjoint a[...] likewise; a holds one extra
// row group (index LT) that absorbs the adjoint writes of non-tape operands.
// The arrays are in LDS (bank-conflict free: 64 lanes touch 64 consecutive
// words) or, for the longest programs, in a global workspace slot.  Each lane
// only touches its own column, so a write needs no synchronisation before
// the same lane reads it back.
template <typename T, int PPT>
struct Tape {
    T *v;
    T *a;
    __device__ inline T val(int i, int p) const { return v[(i * PPT + p) * WAVE + lane_id()]; }
    __device__ inline void set(int i, int p, T x) const { v[(i * PPT + p) * WAVE + lane_id()] = x; }
    __device__ inline T adj(int i, int p) const { return a[(i * PPT + p) * WAVE + lane_id()]; }
    __device__ inline void set_adj(int i, int p, T x) const { a[(i * PPT + p) * WAVE + lane_id()] = x; }
};

// Values of the immediate operands: 0 (unknown leaf), CONST_1..CONST_5,
// PI, E (rounded to T like the reference's literals), NaN (NO_OP).
template <typename T>
__device__ inline T imm_value(int idx) {
    T v = (T)idx;  // 0 (unknown leaf) and 1..5 (CONST_1..CONST_5)
    v = idx == 6 ? (T)3.14159265358979323846 : v;
    v = idx == 7 ? (T)2.71828182845904523536 : v;
    v = idx == IMM_NOOP ? Mth<T>::qnan() : v;
    return v;
}

// Tape of one wave: LT value rows and LT + 1 adjoint rows per point of a
// tile group, 64 elements each.
__host__ __device__ inline size_t tape_elems(int LT, int PPT) { return (size_t)(2 * LT + 1) * PPT * WAVE; }

}  // namespace sr
