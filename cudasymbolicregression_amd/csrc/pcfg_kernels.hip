// Device kernels of the probabilistic context-free grammar: sampling (on the
// project's Philox-4x32-10 stream), infix -> postfix parsing, the postfix
// children table and postfix -> infix rendering.  One lane per expression
// row, 64-lane workgroups; each lane keeps its stack in LDS, laid out
// [depth][lane] so that the lanes of a wave touch neighbouring 16-bit
// slots at the same depth.
//
// The conversions give the tokens and error codes of the host engine
// (pcfg_host.cpp, bit-exact with the reference's src/pcfg_cpu.cpp), whose
// steps they follow:
//   k_infix_to_postfix      src/pcfg_cpu.cpp:261-447 (shunting-yard + arity check)
//   k_postfix_children      src/pcfg_cpu.cpp:1066-1143
//   k_postfix_to_infix      src/pcfg_cpu.cpp:449-635
//   k_sample                src/pcfg_cpu.cpp:10-119 (its derivation loop; the
//                           uniform numbers come from sr_philox.h instead of
//                           std::mt19937, as the reference's CUDA sampler,
//                           src/pcfg_cuda.cu:10-119, draws its own stream too)
// Rows that fail are written as NO_OP (the host engine leaves them as the
// caller filled them, NO_OP in the Python layer).  Rows are at most 128
// tokens (HARD_MAX_LENGTH); the entry points refuse longer ones.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sr_philox.h"

namespace {

constexpr int kLanes = 64;
constexpr int kMaxLen = 128;
constexpr int64_t kNoOp = 0;
constexpr int64_t kNull = -1;

enum : int64_t {
    E_OK = 0,
    E_GEN_MAX_TRIES = 303,
    E_GEN_NO_RULE = 304,
    E_PARSE_QUEUE = 401,
    E_PARSE_STACK = 402,
    E_PARSE_PAREN = 403,
    E_TREE_UNARY = 404,
    E_TREE_BINARY = 405,
    E_TREE_STACK = 406,
    E_TREE_MALFORMED = 407,
    E_PARSE_UNEXPECTED = 408,
    E_CONV_INFIX_LONG = 501,
    E_CONV_UNARY = 505,
    E_CONV_BINARY = 506,
    E_CONV_ROOTS = 508,
    E_CONV_UNEXPECTED = 509,
    E_INTERNAL = 1000,
};

__device__ inline int arity(int64_t op) {
    if (op >= 10 && op <= 31) return 1;
    if (op >= 40 && op <= 44) return 2;
    return 0;
}
__device__ inline bool right_assoc(int64_t op) { return op == 44 || (op >= 10 && op <= 31); }
// unary operators printed as f(x); SQUARE, INV, CUBE, FOURTH, FIFTH as (x)^k
__device__ inline bool functional(int64_t op) { return op >= 10 && op <= 27 && op != 14; }

// Per-lane stack slot `d` of this lane in a [kMaxLen][kLanes] LDS array.
__device__ inline int slot(int d) { return d * kLanes + (int)threadIdx.x; }

__device__ inline void fill_noop(int64_t *row, int64_t n) {
    for (int64_t i = 0; i < n; ++i) row[i] = kNoOp;
}

// ---------------------------------------------------------------- sample
__device__ int64_t sample_row(const int64_t *rhs_ptr, const int64_t *rhs_concat, const int64_t *nt_rule_ptr,
                              const float *cum, int64_t n_symbols, int64_t start_id, int max_length,
                              int64_t max_tries, int64_t *out, sr::PhiloxUniform &unif, int16_t *stk) {
    for (int64_t attempt = 0; attempt < max_tries; ++attempt) {
        int sp = 0, len = 0;
        bool overflow = false;
        stk[slot(sp++)] = (int16_t)start_id;
        while (sp > 0 && !overflow) {
            const int sym = stk[slot(--sp)];
            const int64_t r0 = nt_rule_ptr[sym], r1 = nt_rule_ptr[sym + 1];
            if (r0 == r1) {  // terminal
                if (len >= max_length) {
                    overflow = true;
                    break;
                }
                out[len++] = sym;
                continue;
            }
            const float u = unif();
            int64_t pick = -1;
            for (int64_t r = r0; r < r1; ++r)
                if (u <= cum[r]) {
                    pick = r;
                    break;
                }
            if (pick < 0) return E_GEN_NO_RULE;
            const int64_t s0 = rhs_ptr[pick], s1 = rhs_ptr[pick + 1];
            for (int64_t i = s1 - 1; i >= s0; --i) {
                if (sp >= kMaxLen) {
                    overflow = true;
                    break;
                }
                const int64_t t = rhs_concat[i];
                if (t < 0 || t >= n_symbols) return E_INTERNAL;
                stk[slot(sp++)] = (int16_t)t;
            }
        }
        if (overflow) continue;
        if (sp == 0 && len > 0) {
            for (int i = len; i < max_length; ++i) out[i] = kNoOp;
            return E_OK;
        }
    }
    return E_GEN_MAX_TRIES;
}

__global__ __launch_bounds__(kLanes) void k_sample(const int64_t *rhs_ptr, const int64_t *rhs_concat,
                                                   const int64_t *nt_rule_ptr, const float *cum, int64_t n_symbols,
                                                   int64_t start_id, int max_length, int64_t max_tries,
                                                   int64_t *out, int64_t *errors, uint64_t seed, int64_t B) {
    __shared__ int16_t stk[kMaxLen * kLanes];
    const int64_t b = (int64_t)blockIdx.x * kLanes + threadIdx.x;
    if (b >= B) return;
    sr::PhiloxUniform unif(seed, (uint64_t)b);
    int64_t *row = out + b * max_length;
    const int64_t e = sample_row(rhs_ptr, rhs_concat, nt_rule_ptr, cum, n_symbols, start_id, max_length, max_tries,
                                 row, unif, stk);
    if (e != E_OK) fill_noop(row, max_length);
    errors[b] = e;
}

// ------------------------------------------------------- infix -> postfix
// The output queue is the output row itself: it only ever holds tokens of
// the input (operands and operators), so it never outgrows the row (the
// nq >= M tests cannot fire; they keep every store inside the row).
__device__ int64_t infix_to_postfix_row(const int64_t *prec, int64_t n_prec, const int64_t *in, int64_t *out,
                                        int64_t lp, int64_t rp, int M, int16_t *ops) {
    int len = 0;
    bool ended = false;
    for (int i = 0; i < M; ++i) {
        const int64_t t = in[i];
        if (t < 0 || t >= n_prec) return E_INTERNAL;
        if (t == kNoOp) ended = true;
        else if (ended) return E_PARSE_UNEXPECTED;
        else ++len;
    }
    int nops = 0, nq = 0;
    for (int i = 0; i < len; ++i) {
        const int64_t tok = in[i];
        const int64_t pt = prec[tok];
        if (pt == 0) {
            if (nq >= kMaxLen || nq >= M) return E_PARSE_QUEUE;
            out[nq++] = tok;
        } else if (tok == lp) {
            if (nops >= kMaxLen) return E_PARSE_STACK;
            ops[slot(nops++)] = (int16_t)tok;
        } else if (tok == rp) {
            while (nops > 0 && ops[slot(nops - 1)] != lp) {
                if (nq >= kMaxLen || nq >= M) return E_PARSE_QUEUE;
                out[nq++] = ops[slot(--nops)];
            }
            if (nops == 0) return E_PARSE_PAREN;
            --nops;  // drop '('
        } else {
            const bool ra = right_assoc(tok);
            while (nops > 0 && ops[slot(nops - 1)] != lp) {
                const int64_t top = prec[ops[slot(nops - 1)]];
                if (!((!ra && top >= pt) || (ra && top > pt))) break;
                if (nq >= kMaxLen || nq >= M) return E_PARSE_QUEUE;
                out[nq++] = ops[slot(--nops)];
            }
            if (nops >= kMaxLen) return E_PARSE_STACK;
            ops[slot(nops++)] = (int16_t)tok;
        }
    }
    while (nops > 0) {
        if (ops[slot(nops - 1)] == lp) return E_PARSE_PAREN;
        if (nq >= kMaxLen || nq >= M) return E_PARSE_QUEUE;
        out[nq++] = ops[slot(--nops)];
    }
    if (nq > 0) {  // arity check of the postfix sequence
        int depth = 0;
        for (int i = 0; i < nq; ++i) {
            const int64_t tok = out[i];
            if (prec[tok] == 0) {
                ++depth;
                continue;
            }
            const int a = arity(tok);
            if (depth < a) return a == 1 ? E_TREE_UNARY : (a == 2 ? E_TREE_BINARY : E_INTERNAL);
            depth += 1 - a;
        }
        if (depth != 1) return E_TREE_MALFORMED;
    }
    for (int i = nq; i < M; ++i) out[i] = kNoOp;
    return E_OK;
}

__global__ __launch_bounds__(kLanes) void k_infix_to_postfix(const int64_t *prec, int64_t n_prec, const int64_t *in,
                                                             int64_t *out, int64_t *errors, int64_t lp, int64_t rp,
                                                             int64_t B, int M) {
    __shared__ int16_t ops[kMaxLen * kLanes];
    const int64_t b = (int64_t)blockIdx.x * kLanes + threadIdx.x;
    if (b >= B) return;
    int64_t *row = out + b * M;
    const int64_t e = infix_to_postfix_row(prec, n_prec, in + b * M, row, lp, rp, M, ops);
    if (e != E_OK) fill_noop(row, M);
    errors[b] = e;
}

// ------------------------------------------------------- children table
__device__ int64_t postfix_children_row(const int64_t *in, int64_t *children, int M, int16_t *st) {
    int n = 0, len = 0;
    bool ended = false;
    for (int i = 0; i < M; ++i) {
        const int64_t tok = in[i];
        if (tok == kNoOp) {
            ended = true;
            continue;
        }
        ++len;
        if (ended) return E_PARSE_UNEXPECTED;
        const int a = arity(tok);
        if (a == 1) {
            if (n < 1) return E_TREE_UNARY;
            children[2 * i] = st[slot(--n)];
            children[2 * i + 1] = kNull;
        } else if (a == 2) {
            if (n < 2) return E_TREE_BINARY;
            const int right = st[slot(--n)];
            const int left = st[slot(--n)];
            children[2 * i] = left;
            children[2 * i + 1] = right;
        } else {
            children[2 * i] = kNull;
            children[2 * i + 1] = kNull;
        }
        if (n >= kMaxLen) return E_TREE_STACK;
        st[slot(n++)] = (int16_t)i;
    }
    if (len > 0 && n != 1) return E_TREE_MALFORMED;
    return E_OK;
}

__global__ __launch_bounds__(kLanes) void k_postfix_children(const int64_t *in, int64_t *children, int64_t *errors,
                                                             int64_t B, int M) {
    __shared__ int16_t st[kMaxLen * kLanes];
    const int64_t b = (int64_t)blockIdx.x * kLanes + threadIdx.x;
    if (b >= B) return;
    int64_t *row = children + b * M * 2;
    // NO_OP slots keep -1 (the Python layer's fill)
    for (int i = 0; i < 2 * M; ++i) row[i] = kNull;
    const int64_t e = postfix_children_row(in + b * M, row, M, st);
    if (e != E_OK)
        for (int i = 0; i < 2 * M; ++i) row[i] = kNull;
    errors[b] = e;
}

// ------------------------------------------------------- postfix -> infix
// The host engine builds the string right to left in a buffer and copies it
// to the row's start.  Every operand adds one token and every operator three
// (its parentheses and itself), so a row that converts has
// leaves + 3 * operators tokens, known before the walk: tokens go straight
// to their final place, out[buf_pos - (3 * kMaxLen - total)].  A row whose
// rendering would not fit (total > Mi) is walked without writing, since its
// error code can still be a malformed-tree one.
__device__ int64_t postfix_to_infix_row(const int64_t *in, int64_t *out, int64_t lp, int64_t rp, int M, int Mi,
                                        int16_t *fr) {
    bool started = false;
    int len = 0, balance = 0, roots = 0, nops = 0;
    for (int k = M - 1; k >= 0; --k) {
        const int64_t tok = in[k];
        if (tok == kNoOp) {
            if (started) return E_CONV_UNEXPECTED;
            continue;
        }
        started = true;
        ++len;
        const int a = arity(tok);
        nops += a > 0;
        balance += a - 1;
        if (balance == -1) ++roots;
    }
    if (len == 0) {
        fill_noop(out, Mi);
        return E_OK;
    }
    if (roots > 1) return E_CONV_ROOTS;
    const int total = (len - nops) + 3 * nops;
    const bool write = total <= Mi;
    const int shift = 3 * kMaxLen - total;
    int pos = 3 * kMaxLen - 1;  // next free slot of the host engine's buffer, filled downwards
    auto put = [&](int64_t tok) {
        if (write) out[pos - shift] = tok;
        --pos;
    };
    constexpr int16_t WANT_LEFT = 0x100;
    int nf = 0;
    for (int idx = len - 1; idx >= 0; --idx) {
        const int64_t tok = in[idx];
        const int a = arity(tok);
        if (a == 0) {
            put(tok);
            bool closed = true;
            while (closed && nf > 0) {
                const int16_t f = fr[slot(nf - 1)];
                const int64_t op = f & 0xff;
                if (arity(op) == 1) {
                    put(lp);
                    if (functional(op)) put(op);
                    --nf;
                } else if (!(f & WANT_LEFT)) {
                    put(op);
                    fr[slot(nf - 1)] = (int16_t)(f | WANT_LEFT);
                    closed = false;
                } else {
                    put(lp);
                    --nf;
                }
            }
        } else {
            fr[slot(nf++)] = (int16_t)tok;
            if (a == 1 && !functional(tok)) put(tok);
            put(rp);
        }
    }
    if (nf > 0) return (arity(fr[slot(nf - 1)] & 0xff) == 1) ? E_CONV_UNARY : E_CONV_BINARY;
    if (!write) return E_CONV_INFIX_LONG;
    for (int k = total; k < Mi; ++k) out[k] = kNoOp;
    return E_OK;
}

__global__ __launch_bounds__(kLanes) void k_postfix_to_infix(const int64_t *in, int64_t *out, int64_t *errors,
                                                             int64_t lp, int64_t rp, int64_t B, int M, int Mi) {
    __shared__ int16_t fr[kMaxLen * kLanes];
    const int64_t b = (int64_t)blockIdx.x * kLanes + threadIdx.x;
    if (b >= B) return;
    int64_t *row = out + b * Mi;
    const int64_t e = postfix_to_infix_row(in + b * M, row, lp, rp, M, Mi, fr);
    if (e != E_OK) fill_noop(row, Mi);
    errors[b] = e;
}

inline unsigned grid_of(int64_t B) { return (unsigned)((B + kLanes - 1) / kLanes); }

}  // namespace

extern "C" {

int sr_pcfg_sample_device(const int64_t *rhs_ptr, const int64_t *rhs_concat, const int64_t *nt_rule_ptr,
                          const float *cum_probs, int64_t n_symbols, int64_t start_id, int64_t max_length,
                          int64_t max_tries, int64_t *out, int64_t *errors, uint64_t seed, int64_t B, void *stream) {
    if (B < 0 || max_length < 1 || max_length > kMaxLen || n_symbols < 1 || n_symbols > 32767 || start_id < 0 ||
        start_id >= n_symbols || max_tries < 0 || B > (int64_t)0xffffffffu * kLanes)
        return -1;
    if (B == 0) return 0;
    hipLaunchKernelGGL(k_sample, dim3(grid_of(B)), dim3(kLanes), 0, (hipStream_t)stream, rhs_ptr, rhs_concat,
                       nt_rule_ptr, cum_probs, n_symbols, start_id, (int)max_length, max_tries, out, errors, seed, B);
    return (int)hipGetLastError();
}

int sr_pcfg_infix_to_postfix_device(const int64_t *prec, int64_t n_prec, const int64_t *in, int64_t *out,
                                    int64_t *errors, int64_t lp, int64_t rp, int64_t B, int64_t M, void *stream) {
    if (B < 0 || M < 1 || M > kMaxLen || n_prec < 1 || n_prec > 32767 || B > (int64_t)0xffffffffu * kLanes)
        return -1;
    if (B == 0) return 0;
    hipLaunchKernelGGL(k_infix_to_postfix, dim3(grid_of(B)), dim3(kLanes), 0, (hipStream_t)stream, prec, n_prec, in,
                       out, errors, lp, rp, B, (int)M);
    return (int)hipGetLastError();
}

int sr_pcfg_get_postfix_children_device(const int64_t *in, int64_t *children, int64_t *errors, int64_t B, int64_t M,
                                        void *stream) {
    if (B < 0 || M < 1 || M > kMaxLen || B > (int64_t)0xffffffffu * kLanes) return -1;
    if (B == 0) return 0;
    hipLaunchKernelGGL(k_postfix_children, dim3(grid_of(B)), dim3(kLanes), 0, (hipStream_t)stream, in, children,
                       errors, B, (int)M);
    return (int)hipGetLastError();
}

int sr_pcfg_postfix_to_infix_device(const int64_t *in, int64_t *out, int64_t *errors, int64_t lp, int64_t rp,
                                    int64_t B, int64_t M, int64_t Mi, void *stream) {
    if (B < 0 || M < 1 || M > kMaxLen || Mi < 1 || Mi > 0x7fffffff || B > (int64_t)0xffffffffu * kLanes)
        return -1;
    if (B == 0) return 0;
    hipLaunchKernelGGL(k_postfix_to_infix, dim3(grid_of(B)), dim3(kLanes), 0, (hipStream_t)stream, in, out, errors,
                       lp, rp, B, (int)M, (int)Mi);
    return (int)hipGetLastError();
}

}  // extern "C"
