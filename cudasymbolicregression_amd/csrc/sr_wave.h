// Wave64 helpers and the shared per-expression evaluation core.
//
// Execution model used by every evaluator kernel in this library: one
// wavefront owns one expression.  The expression's packed program (one
// 32-bit word per node, see sr_ops.h) and its constants live in VGPRs with
// lane k holding node k (and k+64 in a second register for M > 64), so that
// decoding node k is a wave-uniform v_readlane into SGPRs and the opcode
// switch is a chain of scalar branches.  Points are mapped to lanes: a tile
// is 64 consecutive points; the per-node values of the tile (the tape) and
// the adjoints live in LDS as [node][lane] arrays, which makes every LDS
// access bank-conflict free.
#pragma once
#include "sr_ops.h"

namespace sr {

constexpr int WAVE = 64;
constexpr int NXREG = 8;   // variables held in registers per point
constexpr int NCREG = 8;   // constants accumulated in per-lane FP64 registers

__device__ inline int lane_id() { return (int)(threadIdx.x & (WAVE - 1)); }

__device__ inline uint32_t rl_u32(uint32_t v, int k) {
    return (uint32_t)__builtin_amdgcn_readlane((int)v, k);
}
__device__ inline float rl(float v, int k) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), k));
}
__device__ inline double rl(double v, int k) {
    int64_t b = __double_as_longlong(v);
    int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffff), k);
    int hi = __builtin_amdgcn_readlane((int)(b >> 32), k);
    return __longlong_as_double(((int64_t)(uint32_t)lo) | ((int64_t)hi << 32));
}

// Values every lane of the wave agrees on, moved to SGPRs so that the
// compiler sees them as uniform: branches and loops that depend on them then
// compile to scalar control flow instead of exec-masked divergent loops.
__device__ inline int32_t uniform_i32(int32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ inline double uniform_f64(double v) {
    int64_t b = __double_as_longlong(v);
    int lo = __builtin_amdgcn_readfirstlane((int)(b & 0xffffffff));
    int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
    return __longlong_as_double(((int64_t)(uint32_t)lo) | ((int64_t)hi << 32));
}

// Next work item of a wave from a global counter.  Every lane issues the
// atomic and only lane 0 adds 1, so there is no lane-dependent branch: with
// an `if (lane == 0)` around the atomic, the optimiser may correlate it with
// another lane-0 branch at the end of the previous item and let lanes 1..63
// re-enter the loop by a shortcut edge that skips the fetch, which turns the
// work loop into an exec-masked loop that re-reads a stale index.  Lane 0's
// ticket is returned to every lane as a wave-uniform value.
__device__ inline int32_t wave_fetch_add(int32_t *counter) {
    const int32_t got = atomicAdd(counter, (threadIdx.x & (64 - 1)) == 0 ? 1 : 0);
    return __builtin_amdgcn_readlane(got, 0);
}

// wave_fetch_add with an increment of n (a chunk of n consecutive tickets).
__device__ inline int32_t wave_fetch_add_n(int32_t *counter, int32_t n) {
    const int32_t got = atomicAdd(counter, (threadIdx.x & (64 - 1)) == 0 ? n : 0);
    return __builtin_amdgcn_readlane(got, 0);
}

// Butterfly sum: every lane computes the same additions in a commuted order,
// so all lanes hold the bit-identical total; it is returned as a uniform.
__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, WAVE);
    return uniform_f64(v);
}
__device__ inline int32_t wave_max_i32(int32_t v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        int32_t o = __shfl_xor(v, off, WAVE);
        v = o > v ? o : v;
    }
    return uniform_i32(v);
}

// DPP reductions.  A row is 16 lanes; quad_perm [1,0,3,2] and [2,3,0,1]
// sum each quad, row_ror:4 and row_ror:8 then leave every lane of a row
// holding its row's total.  These are VALU moves with no LDS round trip.
// All 64 lanes must be active (the library calls them from wave-uniform
// control flow only).
template <int CTRL>
__device__ inline double dpp_f64(double v) {
    const int64_t b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)b, CTRL, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, true);
    return __longlong_as_double(((int64_t)(uint32_t)lo) | ((int64_t)hi << 32));
}
constexpr int DPP_QUAD_SWAP1 = 0xB1;  // quad_perm [1,0,3,2]
constexpr int DPP_QUAD_SWAP2 = 0x4E;  // quad_perm [2,3,0,1]
constexpr int DPP_ROW_ROR4 = 0x124;
constexpr int DPP_ROW_ROR8 = 0x128;

__device__ inline double row_sums(double v) {
    v += dpp_f64<DPP_QUAD_SWAP1>(v);
    v += dpp_f64<DPP_QUAD_SWAP2>(v);
    v += dpp_f64<DPP_ROW_ROR4>(v);
    v += dpp_f64<DPP_ROW_ROR8>(v);
    return v;
}

// Sum over the 64 lanes as a wave-uniform value, in a fixed order
// ((row0 + row1) + (row2 + row3)); `rows` = 1 when only lanes 0..15 can be
// non-zero.
__device__ inline double wave_sum_rows(double v, int rows) {
    v = row_sums(v);
    if (rows == 1) return uniform_f64(rl(v, 0));
    return uniform_f64((rl(v, 0) + rl(v, 16)) + (rl(v, 32) + rl(v, 48)));
}

// Largest per-lane error code among {0, 201, 204} without a shuffle.
__device__ inline int32_t wave_max_err(int32_t e) {
    if (__ballot(e >= 204) != 0ull) return 204;
    if (__ballot(e >= 201) != 0ull) return 201;
    return 0;
}

// Program of one expression held in registers (M <= 128).
template <typename T>
struct Program {
    uint32_t w0, w1;   // packed node words, lane k -> node k / node 64 + k
    T c0, c1;          // constants, same mapping
    uint64_t cmask0, cmask1;  // LEARNABLE_CONSTANT nodes (wave-uniform)
    int L;             // effective length: 1 + last non-NO_OP node (0 if none)
    int root;          // (#non-NO_OP) - 1, clamped at 0
    int32_t verr;      // validation code of this expression (0, 101, 102)

    __device__ inline uint32_t word(int k) const { return k < WAVE ? rl_u32(w0, k) : rl_u32(w1, k - WAVE); }
    __device__ inline T cval(int k) const { return k < WAVE ? rl(c0, k) : rl(c1, k - WAVE); }
    __device__ inline int const_ordinal(int k) const {
        if (k < WAVE) return __popcll(cmask0 & ((k == 0) ? 0ull : (~0ull >> (WAVE - k))));
        int kk = k - WAVE;
        return __popcll(cmask0) + __popcll(cmask1 & ((kk == 0) ? 0ull : (~0ull >> (WAVE - kk))));
    }
};

// Validation of one node against the reference's rule
// (src/evaluation_kernels_cpu.cpp:26-48): every non-null child index must be
// < k (else 101, checked first), and the number of non-null children must
// equal get_arity(op) (else 102).
__device__ inline int32_t validate_node(int64_t op, int64_t c0, int64_t c1, int k) {
    int actual = 0;
    if (c0 != -1) { actual++; if (c0 >= k) return ERR_CHILD_GEQ_PARENT; }
    if (c1 != -1) { actual++; if (c1 >= k) return ERR_CHILD_GEQ_PARENT; }
    int expected = (op >= SIN && op <= FIFTH) ? 1 : ((op >= ADD && op <= POW) ? 2 : 0);
    return actual != expected ? ERR_ARITY : 0;
}

// Loads one expression.  ops/ch/C addresses of node k are given by strides so
// that both the batch-first (B,M) and the backend's node-first (M,B) layouts
// are served: node k of expression b is at ops[b*sb + k*sk], its children at
// ch[(b*sb + k*sk)*2 + i], its constant at C[b*cb + k*ck].
template <typename T>
__device__ inline Program<T> load_program(const int64_t *__restrict__ ops, const int64_t *__restrict__ ch,
                                          const T *__restrict__ C, int64_t b, int64_t sb, int64_t sk,
                                          int64_t cb, int64_t ck, int M, int n_x) {
    Program<T> p;
    const int lane = lane_id();
    uint32_t w[2] = {0u, 0u};
    T cv[2] = {T(0), T(0)};
    int32_t e[2] = {0, 0};
    bool nonnoop[2] = {false, false};
    bool isconst[2] = {false, false};
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        int k = lane + h * WAVE;
        if (k < M) {
            int64_t idx = b * sb + (int64_t)k * sk;
            int64_t op = ops[idx];
            int64_t a = ch[idx * 2], c = ch[idx * 2 + 1];
            w[h] = pack_node(op, a, c, n_x);
            cv[h] = C ? C[b * cb + (int64_t)k * ck] : T(0);
            e[h] = validate_node(op, a, c, k);
            nonnoop[h] = (op != NO_OP);
            isconst[h] = (op == LEARNABLE_CONSTANT);
        }
    }
    p.w0 = w[0];
    p.w1 = w[1];
    p.c0 = cv[0];
    p.c1 = cv[1];
    uint64_t nn0 = __ballot(nonnoop[0]), nn1 = __ballot(nonnoop[1]);
    p.cmask0 = __ballot(isconst[0]);
    p.cmask1 = __ballot(isconst[1]);
    int cnt = __popcll(nn0) + __popcll(nn1);
    p.root = cnt > 0 ? cnt - 1 : 0;
    p.L = nn1 ? (2 * WAVE - __clzll(nn1)) : (nn0 ? WAVE - __clzll(nn0) : 0);
    // code of the first failing node in node order
    uint64_t bad0 = __ballot(e[0] != 0), bad1 = __ballot(e[1] != 0);
    p.verr = bad0 ? __builtin_amdgcn_readlane(e[0], __ffsll((long long)bad0) - 1)
                  : (bad1 ? __builtin_amdgcn_readlane(e[1], __ffsll((long long)bad1) - 1) : 0);
    return p;
}

// Point-tile context: variable values of this lane's point.
template <typename T>
struct PointCtx {
    T x[NXREG];
    const T *xrow;   // row pointer for n_x > NXREG
    int n_x;
    __device__ inline T var(int v) const {
        if (n_x <= NXREG) {
            T r = T(0);
#pragma unroll
            for (int q = 0; q < NXREG; ++q) r = (q == v) ? x[q] : r;
            return r;
        }
        return xrow[v];
    }
};

template <typename T>
__device__ inline PointCtx<T> load_point(const T *__restrict__ X, int64_t n, bool active, int n_x) {
    PointCtx<T> pc;
    pc.n_x = n_x;
    pc.xrow = X + (active ? n : 0) * (int64_t)n_x;
#pragma unroll
    for (int q = 0; q < NXREG; ++q) pc.x[q] = (active && q < n_x) ? pc.xrow[q] : T(0);
    return pc;
}

// Forward sweep of one tile over nodes [0, L): tape[k*64 + lane].
// Forward sweep over nodes [0, L) on any tape: at(k) is this lane's element
// of node k's row (an LDS tape, or rows in global memory).
template <typename T, typename At>
__device__ inline void forward_tile_at(const Program<T> &p, const PointCtx<T> &pc, At at, int L) {
    for (int k = 0; k < L; ++k) {
        uint32_t w = p.word(k);
        int op = node_op(w);
        T r;
        if (op == NO_OP) {
            r = Mth<T>::qnan();
        } else if (op == LEARNABLE_CONSTANT) {
            r = p.cval(k);
        } else if (op == UNKNOWN_LEAF) {
            r = T(0);
        } else if (op >= VAR_START_ID) {
            r = pc.var(op - VAR_START_ID);
        } else {
            int ar = arity_of(op);
            T a0 = ar >= 1 ? at(node_c0(w)) : T(0);
            T a1 = ar == 2 ? at(node_c1(w)) : T(0);
            r = fwd_node<T>(op, a0, a1);
        }
        at(k) = r;
    }
}

template <typename T>
__device__ inline void forward_tile(const Program<T> &p, const PointCtx<T> &pc, T *tape, int L) {
    const int lane = lane_id();
    forward_tile_at<T>(p, pc, [&](int k) -> T & { return tape[k * WAVE + lane]; }, L);
}

// Per-lane accumulators of one expression over all of its tiles.
template <typename T>
struct GradAcc {
    double cacc[NCREG];  // constants with ordinal < NCREG (FP64 per lane)
    T gxl[NXREG];        // variable adjoints of the current tile
};

// Reverse sweep of one tile over nodes [0, L) with adjoints preloaded.
// `val(k)` returns this lane's forward value of node k and `adj(k)` a
// reference to its adjoint (LDS tape, a global tape, or the caller's
// grad_cache).  Constant adjoints go to acc.cacc (first NCREG constants) or,
// beyond that, to a wave-uniform FP64 LDS slot cextra[k].  Variable adjoints
// go to acc.gxl[v] when n_x <= NXREG; with more variables the caller sums
// them after the sweep with var_adjoints (same values, same order).
// Inactive lanes (points past N) read but never write and contribute
// nothing.  Returns this lane's error code (0, 201 or 204; the larger one if
// both occur).
template <typename T, bool WANT_GX, typename Val, typename Adj>
__device__ inline int32_t backward_tile(const Program<T> &p, Val val, Adj adj, int L, bool active,
                                        GradAcc<T> &acc, double *cextra, int n_x) {
    const int lane = lane_id();
    int32_t err = 0;
    for (int k = L - 1; k >= 0; --k) {
        uint32_t w = p.word(k);
        int op = node_op(w);
        T g = adj(k);
        if (op == NO_OP) {
            if (active && g != T(0)) err = err > ERR_GRAD_NO_OP ? err : ERR_GRAD_NO_OP;
            continue;
        }
        if (op == LEARNABLE_CONSTANT) {
            int j = p.const_ordinal(k);
            double gv = (active && g != T(0)) ? (double)g : 0.0;
            if (j < NCREG) {
#pragma unroll
                for (int q = 0; q < NCREG; ++q)
                    if (q == j) acc.cacc[q] += gv;
            } else {
                double s = wave_sum(gv);
                if (lane == 0) cextra[k] += s;
            }
            continue;
        }
        if (op == UNKNOWN_LEAF || op < SIN) continue;  // hard constants and 0-valued leaves
        if (op >= VAR_START_ID) {
            if (WANT_GX && active && g != T(0) && n_x <= NXREG) {
                int v = op - VAR_START_ID;
#pragma unroll
                for (int q = 0; q < NXREG; ++q)
                    if (q == v) acc.gxl[q] += g;
            }
            continue;
        }
        int c0 = node_c0(w), c1 = node_c1(w);
        int ar = arity_of(op);
        T a0 = val(c0);
        T a1 = ar == 2 ? val(c1) : T(0);
        T go0 = T(0), go1 = T(0);
        bool add0, add1;
        int32_t e = bwd_node<T>(op, g, a0, a1, go0, add0, go1, add1);
        if (active && e > err) err = e;
        if (active && add0) adj(c0) += go0;
        if (active && add1) adj(c1) += go1;
    }
    return err;
}

// Adjoint sums of the variables vbase .. vbase + NXREG - 1 at this lane's
// point, from the final node adjoints adj(k) after a reverse sweep: the
// variable nodes' adjoints added in reverse node order, as backward_tile adds
// them during the sweep (a variable node's adjoint is final when the sweep
// reaches it).  Used when n_x > NXREG, one chunk of variables at a time.
template <typename T, typename Adj>
__device__ inline void var_adjoints(const Program<T> &p, Adj adj, int L, bool active, int vbase, T (&gxl)[NXREG]) {
#pragma unroll
    for (int q = 0; q < NXREG; ++q) gxl[q] = T(0);
    for (int k = L - 1; k >= 0; --k) {
        const int op = node_op(p.word(k));
        if (op < VAR_START_ID || op == UNKNOWN_LEAF) continue;
        const int v = op - VAR_START_ID - vbase;
        if (v < 0 || v >= NXREG) continue;
        const T g = adj(k);
        if (active && g != T(0))
#pragma unroll
            for (int q = 0; q < NXREG; ++q)
                if (q == v) gxl[q] += g;
    }
}

// Wave-reduces the constant accumulators and returns, in lane k, the FP64
// gradient of node k (0 for non-constant nodes); `hi` selects nodes 64..127.
template <typename T>
__device__ inline void finish_constants(const Program<T> &p, const GradAcc<T> &acc, const double *cextra,
                                        double &out_lo, double &out_hi) {
    const int lane = lane_id();
    out_lo = 0.0;
    out_hi = 0.0;
    int j = 0;
#pragma unroll 1
    for (int h = 0; h < 2; ++h) {
        uint64_t m = h == 0 ? p.cmask0 : p.cmask1;
        while (m) {
            int bit = __ffsll((long long)m) - 1;
            m &= m - 1;
            int k = bit + h * WAVE;
            double s;
            if (j < NCREG) {
                double v = 0.0;
#pragma unroll
                for (int q = 0; q < NCREG; ++q)
                    if (q == j) v = acc.cacc[q];
                s = wave_sum(v);
            } else {
                s = cextra[k];
            }
            if (lane == bit) {
                if (h == 0) out_lo = s; else out_hi = s;
            }
            ++j;
        }
    }
}

// s + src[r0 * stride] + src[(r0 + 1) * stride] + ... + src[(r1 - 1) * stride],
// added in FP64 in row order; the loads go out 16 at a time, so a thread
// pays one memory latency per 16 rows instead of one per row.
template <typename T>
__device__ inline double sum_rows_in_order(const T *__restrict__ src, int64_t r0, int64_t r1, int64_t stride,
                                           double s) {
    constexpr int U = 16;
    int64_t r = r0;
    for (; r + U <= r1; r += U) {
        T v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = src[(r + u) * stride];
#pragma unroll
        for (int u = 0; u < U; ++u) s += (double)v[u];
    }
    for (; r < r1; ++r) s += (double)src[r * stride];
    return s;
}

}  // namespace sr
