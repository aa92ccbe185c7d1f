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
__device__ inline CProg<T> load_cprog(const int64_t *ops, const int64_t *ch, const T *C, int64_t b, int M, int n_x,
                                      int64_t cb = -1) {
    // program of row b; constants of row cb (b when cb < 0)
    Program<T> p = load_program<T>(ops + b * M, ch + b * M * 2, C ? C + (cb < 0 ? b : cb) * M : nullptr, 0, M, 1,
                                   M, 1, M, n_x);
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
constexpr int MAX_NX_CODE = 2047;

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

// Selects rather than stores under a branch: a store to either field under
// a condition becomes a store through a computed address, which puts the
// struct in private (scratch) memory.
__device__ inline void place_word(Code &c, int i, uint32_t v) {
    const int lane = lane_id();
    const bool hi = i >= WAVE;
    const bool mine = lane == (hi ? i - WAVE : i);
    c.iw0 = (mine && !hi) ? v : c.iw0;
    c.iw1 = (mine && hi) ? v : c.iw1;
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

// Tape of one wave: row r at point p of a tile group is
// v[(r*PPT + p)*64 + lane].  Rows 0..LT-1 hold instruction values and
// adjoints (below), rows LT..LT+NXR_ROWS-1 the tile group's first variables.
// The array is in LDS (bank-conflict free: 64 lanes touch 64 consecutive
// words) or, for the longest programs, in a global workspace slot.  Each
// lane only touches its own column, so a write needs no synchronisation
// before the same lane reads it back.
constexpr int NXR_ROWS = 4;
template <typename T, int PPT>
struct Tape {
    T *v;
    __device__ inline T get(int r, int p) const { return v[(r * PPT + p) * WAVE + lane_id()]; }
    __device__ inline void put(int r, int p, T x) const { v[(r * PPT + p) * WAVE + lane_id()] = x; }
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

// Tape of one wave: LT instruction rows and NXR_ROWS variable rows per point
// of a tile group, 64 elements each.
__host__ __device__ inline size_t tape_elems(int LT, int PPT) { return (size_t)(LT + NXR_ROWS) * PPT * WAVE; }

// ------------------------------------------------------------------ plans
// The instruction list re-encoded for evaluation (one fetch word and one
// scatter word per instruction, lane i and 64 + i):
//   * an operand is a tape row (an instruction, or a variable v < NXR_ROWS
//     in row LT + v), a wave-uniform scalar (a constant or an immediate,
//     resolved per evaluation into sv0 / sv1, see resolve_scalars), or a
//     variable read from X (v >= NXR_ROWS);
//   * adjoints overwrite values in place.  In a tree every instruction below
//     the root has one parent, and the parent's backward step is both the
//     last use of the child's value and the only write of its adjoint.  An
//     instruction with several parents, or with none below the root, keeps
//     its value row and gets an extra adjoint row (NI, NI + 1, ...) that is
//     zeroed before each tile group and accumulated into, in the same order
//     as before;
//   * a constant used once has its adjoint written into a row (the row of
//     the instruction that uses it, dead by then, or an extra row for the
//     second constant operand of one instruction; crow0 / crow1 name it per
//     constant), and the evaluators add those rows into their FP64
//     per-constant sums after each tile group.  A constant used more than
//     once is reduced over the wave at every use (PM_CONST).
// The bucket of a program is chosen by its row count NI + extra rows
// (count_rows_wave).  With n operator nodes and l leaves (n + l <= M), the
// rows are at most (n + 1) + n + l / 2 <= 2 M - 1.
// Fetch word: op 0-5 | binary 6 | scalar0 7 | row0 8-16 | scalar1 17 |
// row1 18-26 | xg0 27 | xg1 28 (xg: variable read from X).
// Scatter word: adjoint row of the instruction 0-7 | mode0 8-10 |
// target0 11-18 | mode1 19-21 | target1 22-29.
// Modes 1-3 write a row (the value 0 + g, or the row's previous value + g
// for PM_ACC), so one range test selects them.
constexpr uint32_t PM_NONE = 0, PM_WRITE = 1, PM_CROW = 2, PM_ACC = 3, PM_VAR = 4, PM_CONST = 5, PM_NOOP = 6;
constexpr uint32_t NO_CROW = 255;

__host__ __device__ inline int pf_op(uint32_t f) { return (int)(f & 63u); }
__host__ __device__ inline bool pf_bin(uint32_t f) { return (f >> 6) & 1u; }
__host__ __device__ inline bool pf_sc(uint32_t f, int k) { return (f >> (k ? 17 : 7)) & 1u; }
__host__ __device__ inline int pf_row(uint32_t f, int k) { return (int)((f >> (k ? 18 : 8)) & 511u); }
__host__ __device__ inline bool pf_xg(uint32_t f, int k) { return (f >> (k ? 28 : 27)) & 1u; }
__host__ __device__ inline int ps_grow(uint32_t s) { return (int)(s & 255u); }
__host__ __device__ inline uint32_t ps_mode(uint32_t s, int k) { return (s >> (k ? 19 : 8)) & 7u; }
__host__ __device__ inline int ps_tgt(uint32_t s, int k) { return (int)((s >> (k ? 22 : 11)) & 255u); }

struct Plan {
    uint32_t f0, f1;      // fetch words
    uint32_t s0, s1;      // scatter words
    uint32_t crow0, crow1;  // lane j: adjoint row of constant j (and 64 + j), or NO_CROW
    int nextra;           // extra rows (NI .. NI + nextra - 1)
    template <int LT>
    __device__ inline uint32_t fw(int i) const {
        if constexpr (LT <= WAVE) return rl_u32(f0, i);
        else return i < WAVE ? rl_u32(f0, i) : rl_u32(f1, i - WAVE);
    }
    template <int LT>
    __device__ inline uint32_t sw(int i) const {
        if constexpr (LT <= WAVE) return rl_u32(s0, i);
        else return i < WAVE ? rl_u32(s0, i) : rl_u32(s1, i - WAVE);
    }
    __device__ inline int crow(int j) const { return (int)(j < WAVE ? rl_u32(crow0, j) : rl_u32(crow1, j - WAVE)); }
};

// Uses of every node (lane k, 64 + k) by the operator nodes om0 / om1 of the
// node words w0 / w1, plus the copy instruction of a leaf root; the extra
// instruction rows (operator nodes below the root used other than once), and
// the constants used exactly once (cs0 / cs1).
__device__ inline void node_uses(uint32_t w0, uint32_t w1, uint64_t om0, uint64_t om1, int root, int M, uint64_t &em0,
                                 uint64_t &em1, uint64_t &cs0, uint64_t &cs1) {
    const int lane = lane_id();
    int rc0 = lane == root ? 1 : 0, rc1 = lane + WAVE == root ? 1 : 0;  // the copy of a leaf root
    auto count = [&](uint32_t w) {
        const int c0 = node_c0(w), c1 = arity_of(node_op(w)) == 2 ? node_c1(w) : -1;
        rc0 += (c0 == lane ? 1 : 0) + (c1 == lane ? 1 : 0);
        rc1 += (c0 == lane + WAVE ? 1 : 0) + (c1 == lane + WAVE ? 1 : 0);
    };
    for (uint64_t m = om0; m; m &= m - 1) count(rl_u32(w0, __ffsll((long long)m) - 1));
    for (uint64_t m = om1; m; m &= m - 1) count(rl_u32(w1, __ffsll((long long)m) - 1));
    const bool op0 = (om0 >> lane) & 1ull, op1 = (om1 >> lane) & 1ull;
    // an operator root is not a copy's operand: its count is its parents'
    em0 = __ballot(op0 && lane < root && rc0 != 1);
    em1 = __ballot(op1 && lane + WAVE < root && rc1 != 1);
    cs0 = __ballot(lane < M && node_op(w0) == LEARNABLE_CONSTANT && rc0 == 1);
    cs1 = __ballot(lane + WAVE < M && node_op(w1) == LEARNABLE_CONSTANT && rc1 == 1);
}

// Binary operator nodes whose two operands are constants used once: each
// needs one extra row for its second constant's adjoint.
__device__ inline void second_const_rows(uint32_t w0, uint32_t w1, uint64_t om0, uint64_t om1, uint64_t cs0,
                                         uint64_t cs1, uint64_t &dm0, uint64_t &dm1) {
    const int lane = lane_id();
    auto is_cs = [&](int c) -> bool { return c < WAVE ? ((cs0 >> c) & 1ull) : (c < 2 * WAVE && ((cs1 >> (c - WAVE)) & 1ull)); };
    auto both = [&](uint32_t w) -> bool {
        return arity_of(node_op(w)) == 2 && is_cs(node_c0(w)) && is_cs(node_c1(w));
    };
    dm0 = __ballot(((om0 >> lane) & 1ull) && both(w0));
    dm1 = __ballot(((om1 >> lane) & 1ull) && both(w1));
}

// Tape rows a batch-first program row needs (NI + extra rows), under the
// same rules as compile_code and compile_plan; one wave per row, node k in
// lane k (and 64 + k), reads coalesced.  Returns wave-uniform values.
__device__ inline int count_rows_wave(const int64_t *ops_row, const int64_t *ch_row, int M, int n_x, int &nc) {
    const int lane = lane_id();
    uint32_t w[2] = {0u, 0u};
    bool nn[2] = {false, false}, isc[2] = {false, false};
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int k = lane + h * WAVE;
        if (k < M) {
            const int64_t op = ops_row[k];
            w[h] = pack_node(op, ch_row[2 * k], ch_row[2 * k + 1], n_x);
            nn[h] = op != NO_OP;
            isc[h] = op == LEARNABLE_CONSTANT;
        }
    }
    const uint64_t nn0 = __ballot(nn[0]), nn1 = __ballot(nn[1]);
    nc = __popcll(__ballot(isc[0])) + __popcll(__ballot(isc[1]));
    const int cnt = __popcll(nn0) + __popcll(nn1);
    const int root = cnt > 0 ? cnt - 1 : 0;
    const uint64_t om0 = __ballot(lane < M && lane <= root && arity_of(node_op(w[0])) > 0);
    const uint64_t om1 = __ballot(lane + WAVE < M && lane + WAVE <= root && arity_of(node_op(w[1])) > 0);
    const bool root_is_op = root < WAVE ? ((om0 >> root) & 1ull) : ((om1 >> (root - WAVE)) & 1ull);
    const int ni = __popcll(om0) + __popcll(om1) + (root_is_op ? 0 : 1);
    uint64_t em0, em1, cs0, cs1, dm0, dm1;
    node_uses(w[0], w[1], om0, om1, root, M, em0, em1, cs0, cs1);
    second_const_rows(w[0], w[1], om0, om1, cs0, cs1, dm0, dm1);
    return ni + __popcll(em0) + __popcll(em1) + __popcll(dm0) + __popcll(dm1);
}

// Extra tape rows of a loaded program (the rows beyond its instructions, as
// compile_plan assigns them); NI + extra_rows is the program's row count.
template <typename T>
__device__ inline int extra_rows(const CProg<T> &p, int M) {
    const int lane = lane_id();
    const int root = p.root;
    const uint64_t om0 = __ballot(lane < M && lane <= root && arity_of(node_op(p.w0)) > 0);
    const uint64_t om1 = __ballot(lane + WAVE < M && lane + WAVE <= root && arity_of(node_op(p.w1)) > 0);
    uint64_t em0, em1, cs0, cs1, dm0, dm1;
    node_uses(p.w0, p.w1, om0, om1, root, M, em0, em1, cs0, cs1);
    second_const_rows(p.w0, p.w1, om0, om1, cs0, cs1, dm0, dm1);
    return __popcll(em0) + __popcll(em1) + __popcll(dm0) + __popcll(dm1);
}

// Operand record of a node (used by its parent): scalar 0 | row 1-9 |
// xg 10 | scatter mode 11-13 | scatter target 14-21 | constant used once 22.
// var_adj_row0 >= 0: a variable v < nvr accumulates its adjoint in row
// var_adj_row0 + v (PM_ACC) instead of PM_VAR.
__device__ inline uint32_t operand_rec(uint32_t w, bool is_op, uint32_t ix, bool extra, uint32_t erow, bool once,
                                       int var_row0, int nvr, int var_adj_row0) {
    if (is_op) return (ix << 1) | ((extra ? PM_ACC : PM_WRITE) << 11) | ((extra ? erow : ix) << 14);
    const uint32_t d = operand_desc(w, 0u, false);
    const int kind = desc_kind(d);
    const uint32_t idx = (uint32_t)desc_idx(d);
    if (kind == K_VAR) {
        const uint32_t t = (idx > 255u ? 255u : idx) << 14;
        if ((int)idx < nvr && var_adj_row0 >= 0)
            return (((uint32_t)var_row0 + idx) << 1) | (PM_ACC << 11) | (((uint32_t)var_adj_row0 + idx) << 14);
        return (int)idx < nvr ? (((uint32_t)var_row0 + idx) << 1) | (PM_VAR << 11) | t : (1u << 10) | (PM_VAR << 11) | t;
    }
    if (kind == K_CONST) return 1u | (PM_CONST << 11) | (idx << 14) | ((once ? 1u : 0u) << 22);
    return 1u | ((idx == (uint32_t)IMM_NOOP ? PM_NOOP : PM_NONE) << 11);
}

// Builds the plan of a compiled program (c = compile_code(p, M)).  Variables
// v < nvr are tape rows var_row0 + v; the others are read from X.  With
// var_adj_row0 >= 0 the adjoints of the variables v < nvr accumulate in rows
// var_adj_row0 + v (at most 255), which the caller zeroes per tile group.
template <typename T>
__device__ inline Plan compile_plan(const CProg<T> &p, int M, const Code &c, int var_row0, int nvr,
                                    int var_adj_row0 = -1) {
    const int lane = lane_id();
    const int root = p.root;
    const int NI = c.NI;
    const uint32_t w0 = p.w0, w1 = p.w1;
    const bool op0 = lane < M && lane <= root && arity_of(node_op(w0)) > 0;
    const bool op1 = lane + WAVE < M && lane + WAVE <= root && arity_of(node_op(w1)) > 0;
    const uint64_t om0 = __ballot(op0), om1 = __ballot(op1);
    const uint64_t below = lane == 0 ? 0ull : (~0ull >> (WAVE - lane));
    const uint32_t ix0 = (uint32_t)__popcll(om0 & below);
    const uint32_t ix1 = (uint32_t)(__popcll(om0) + __popcll(om1 & below));
    uint64_t em0, em1, cs0, cs1, dm0, dm1;
    node_uses(w0, w1, om0, om1, root, M, em0, em1, cs0, cs1);
    second_const_rows(w0, w1, om0, om1, cs0, cs1, dm0, dm1);
    const bool ex0 = (em0 >> lane) & 1ull, ex1 = (em1 >> lane) & 1ull;
    const int ne = __popcll(em0) + __popcll(em1);
    const uint32_t er0 = (uint32_t)(NI + __popcll(em0 & below));
    const uint32_t er1 = (uint32_t)(NI + __popcll(em0) + __popcll(em1 & below));
    // second-constant rows follow the extra instruction rows
    const uint32_t dr0 = (uint32_t)(NI + ne + __popcll(dm0 & below));
    const uint32_t dr1 = (uint32_t)(NI + ne + __popcll(dm0) + __popcll(dm1 & below));
    const uint32_t r0 = operand_rec(w0, op0, ix0, ex0, er0, (cs0 >> lane) & 1ull, var_row0, nvr, var_adj_row0);
    const uint32_t r1 = operand_rec(w1, op1, ix1, ex1, er1, (cs1 >> lane) & 1ull, var_row0, nvr, var_adj_row0);
    auto child_rec = [&](int ch) -> uint32_t {  // every lane takes part in both shuffles
        const int cc = ch & 127;
        const uint32_t lo = (uint32_t)__shfl((int)r0, cc & (WAVE - 1), WAVE);
        const uint32_t hi = (uint32_t)__shfl((int)r1, cc & (WAVE - 1), WAVE);
        return cc < WAVE ? lo : hi;
    };
    auto fetch_word = [](uint32_t op, bool bin, uint32_t a, uint32_t b) -> uint32_t {
        b = bin ? b : 0u;
        return op | ((bin ? 1u : 0u) << 6) | ((a & 1u) << 7) | (((a >> 1) & 511u) << 8) | ((b & 1u) << 17) |
               (((b >> 1) & 511u) << 18) | (((a >> 10) & 1u) << 27) | (((b >> 10) & 1u) << 28);
    };
    // a constant used once writes its adjoint to crow_a (operand 0) or crow_b
    auto scatter_mode = [](uint32_t a, uint32_t crow) -> uint32_t {
        const bool once = (a >> 22) & 1u;
        return once ? (PM_CROW | (crow << 3)) : (((a >> 11) & 7u) | (((a >> 14) & 255u) << 3));
    };
    auto scatter_word = [&](uint32_t grow, bool bin, uint32_t a, uint32_t b, uint32_t crow_a,
                            uint32_t crow_b) -> uint32_t {
        const uint32_t ma = scatter_mode(a, crow_a), mb = bin ? scatter_mode(b, crow_b) : 0u;
        return grow | ((ma & 7u) << 8) | ((ma >> 3) << 11) | ((mb & 7u) << 19) | ((mb >> 3) << 22);
    };
    const uint32_t a00 = child_rec(node_c0(w0)), a01 = child_rec(node_c1(w0));
    const uint32_t a10 = child_rec(node_c0(w1)), a11 = child_rec(node_c1(w1));
    const bool b0 = arity_of(node_op(w0)) == 2, b1 = arity_of(node_op(w1)) == 2;
    const bool d0 = (dm0 >> lane) & 1ull, d1 = (dm1 >> lane) & 1ull;
    // rows of the constant operands: the instruction's own row, and for the
    // second of two constants its second-constant row
    const uint32_t ca0 = ix0, cb0 = d0 ? dr0 : ix0, ca1 = ix1, cb1 = d1 ? dr1 : ix1;
    const uint32_t F0 = fetch_word((uint32_t)node_op(w0) & 63u, b0, a00, a01);
    const uint32_t F1 = fetch_word((uint32_t)node_op(w1) & 63u, b1, a10, a11);
    const uint32_t S0 = scatter_word(ex0 ? er0 : ix0, b0, a00, a01, ca0, cb0);
    const uint32_t S1 = scatter_word(ex1 ? er1 : ix1, b1, a10, a11, ca1, cb1);
    Plan pl;
    pl.f0 = pl.f1 = pl.s0 = pl.s1 = 0u;
    pl.crow0 = pl.crow1 = NO_CROW;
    pl.nextra = ne + __popcll(dm0) + __popcll(dm1);
    // (selects, as in place_word, keep the plan in registers)
    auto place = [&](int i, uint32_t f, uint32_t s) {
        const bool hi = i >= WAVE;
        const bool mine = lane == (hi ? i - WAVE : i);
        pl.f0 = (mine && !hi) ? f : pl.f0;
        pl.s0 = (mine && !hi) ? s : pl.s0;
        pl.f1 = (mine && hi) ? f : pl.f1;
        pl.s1 = (mine && hi) ? s : pl.s1;
    };
    // the constant ordinal of an operand record, and its row, into lane o
    auto place_crow = [&](uint32_t rec, uint32_t row) {
        const bool once = (rec >> 22) & 1u;
        const int o = (int)((rec >> 14) & 255u);
        const bool hi = o >= WAVE;
        const bool mine = once && lane == (hi ? o - WAVE : o);
        pl.crow0 = (mine && !hi) ? row : pl.crow0;
        pl.crow1 = (mine && hi) ? row : pl.crow1;
    };
    int i = 0;
    for (uint64_t m = om0; m; m &= m - 1, ++i) {
        const int k = __ffsll((long long)m) - 1;
        place(i, rl_u32(F0, k), rl_u32(S0, k));
        place_crow(rl_u32(a00, k), rl_u32(ca0, k));
        if (rl_u32((uint32_t)b0, k)) place_crow(rl_u32(a01, k), rl_u32(cb0, k));
    }
    for (uint64_t m = om1; m; m &= m - 1, ++i) {
        const int k = __ffsll((long long)m) - 1;
        place(i, rl_u32(F1, k), rl_u32(S1, k));
        place_crow(rl_u32(a10, k), rl_u32(ca1, k));
        if (rl_u32((uint32_t)b1, k)) place_crow(rl_u32(a11, k), rl_u32(cb1, k));
    }
    if (i < NI) {  // the root is a leaf: one copy instruction
        const uint32_t rr = root < WAVE ? rl_u32(r0, root) : rl_u32(r1, root - WAVE);
        place(i, fetch_word((uint32_t)OP_COPY, false, rr, 0u), scatter_word((uint32_t)i, false, rr, 0u, (uint32_t)i, 0u));
        place_crow(rr, (uint32_t)i);
    }
    return pl;
}

// Scalar operand values of every instruction (lane i, and 64 + i when LT >
// 64): the constant (compact slot) or immediate each operand names, 0 for
// row operands.  Recomputed whenever the constants change.
template <typename T, int SL, int LT>
__device__ inline void resolve_scalars(const Code &c, const SVec<T, SL> &xc, T (&sv0)[LT > WAVE ? 2 : 1],
                                       T (&sv1)[LT > WAVE ? 2 : 1]) {
    constexpr int H = LT > WAVE ? 2 : 1;
#pragma unroll
    for (int h = 0; h < H; ++h) {
        const uint32_t w = h == 0 ? c.iw0 : c.iw1;
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const uint32_t d = k == 0 ? instr_d0(w) : instr_d1(w);
            const int kind = desc_kind(d), idx = desc_idx(d);
            const int j = kind == K_CONST ? idx : 0;
            T cv = __shfl(xc.v[0], j & (WAVE - 1), WAVE);
            if constexpr (SL > 1) {
                const T hv = __shfl(xc.v[SL - 1], j & (WAVE - 1), WAVE);
                cv = j < WAVE ? cv : hv;
            }
            const T v = kind == K_CONST ? cv : (kind == K_IMM ? imm_value<T>(idx) : T(0));
            if (k == 0) sv0[h] = v;
            else sv1[h] = v;
        }
    }
}

}  // namespace sr
