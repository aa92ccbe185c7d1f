// Host engine of the probabilistic context-free grammar: sampling, infix
// parsing to postfix/prefix, conversions back to infix, and the parent /
// children tables.  One OpenMP iteration per expression row.
//
// The per-row algorithms must give bit-identical tokens and error codes to the
// reference's CPU engine (src/pcfg_cpu.cpp), so their steps follow it:
//   sample_row            src/pcfg_cpu.cpp:10-119  (std::mt19937 seeded per row,
//                         uniform_real_distribution<float>, restart on overflow)
//   infix_to_postfix_row  src/pcfg_cpu.cpp:261-447 (shunting-yard + arity check)
//   infix_to_prefix_row   src/pcfg_cpu.cpp:121-259 (right-to-left shunting-yard)
//   postfix_to_infix_row  src/pcfg_cpu.cpp:449-635
//   prefix_to_infix_row   src/pcfg_cpu.cpp:637-810
//   prefix_to_postfix_row src/pcfg_cpu.cpp:812-900
//   postfix_to_prefix_row src/pcfg_cpu.cpp:902-993
//   prefix_parent_row     src/pcfg_cpu.cpp:995-1064
//   postfix_children_row  src/pcfg_cpu.cpp:1066-1143
// Tables are plain row-major int64 arrays owned by the caller; output arrays
// must be pre-filled by the caller exactly like the reference's host code
// does (NO_OP for token outputs, -1 for parent/children tables), because a
// failing row leaves them untouched.
#include <algorithm>
#include <cstdint>
#include <numeric>
#include <random>
#include <vector>

#include "sr_philox.h"

namespace {

constexpr int kMaxLen = 128;  // hard cap on stacks and queues (HARD_MAX_LENGTH)
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
    E_CONV_POSTFIX_LONG = 502,
    E_CONV_PREFIX_LONG = 503,
    E_CONV_STACK = 504,
    E_CONV_UNARY = 505,
    E_CONV_BINARY = 506,
    E_CONV_ROOTS = 508,
    E_CONV_UNEXPECTED = 509,
    E_INTERNAL = 1000,
};

inline int arity(int64_t op) {
    if (op >= 10 && op <= 31) return 1;
    if (op >= 40 && op <= 44) return 2;
    return 0;
}
// ADD SUB MUL DIV are left-associative; POW and every unary operator are
// right-associative; anything else counts as left-associative.
inline bool right_assoc(int64_t op) { return op == 44 || (op >= 10 && op <= 31); }
inline bool left_assoc(int64_t op) { return !right_assoc(op); }
// Unary operators printed as f(x); SQUARE, INV, CUBE, FOURTH, FIFTH print as (x)^k.
inline bool functional(int64_t op) {
    return (op >= 10 && op <= 27 && op != 14);
}

// ------------------------------------------------------------------ sample
// Leftmost derivation with an explicit symbol stack; `unif` gives the
// uniform number of each rule choice.
template <class Uniform>
int64_t sample_row_with(const int64_t *rhs_ptr, const int64_t *rhs_concat, const int64_t *nt_rule_ptr,
                        const float *cum, int64_t start_id, int64_t max_length, int64_t max_tries, int64_t *out,
                        Uniform &unif) {
    for (int64_t attempt = 0; attempt < max_tries; ++attempt) {
        int32_t stack[kMaxLen];
        int32_t sp = 0;
        int32_t len = 0;
        bool overflow = false;
        stack[sp++] = (int32_t)start_id;
        while (sp > 0 && !overflow) {
            const int32_t sym = stack[--sp];
            const int32_t r0 = (int32_t)nt_rule_ptr[sym], r1 = (int32_t)nt_rule_ptr[sym + 1];
            if (r0 == r1) {  // terminal
                if (len >= max_length) { overflow = true; break; }
                out[len++] = sym;
                continue;
            }
            const float u = unif();
            int32_t pick = -1;
            for (int32_t r = r0; r < r1; ++r)
                if (u <= cum[r]) { pick = r; break; }
            if (pick < 0) return E_GEN_NO_RULE;
            const int32_t s0 = (int32_t)rhs_ptr[pick], s1 = (int32_t)rhs_ptr[pick + 1];
            for (int32_t i = s1 - 1; i >= s0; --i) {
                if (sp >= kMaxLen) { overflow = true; break; }
                stack[sp++] = (int32_t)rhs_concat[i];
            }
        }
        if (overflow) continue;
        if (sp == 0 && len > 0) {
            for (int64_t i = len; i < max_length; ++i) out[i] = kNoOp;
            return E_OK;
        }
    }
    return E_GEN_MAX_TRIES;
}

// The reference's stream: std::mt19937 seeded per row with
// uniform_real_distribution<float>.
int64_t sample_row(const int64_t *rhs_ptr, const int64_t *rhs_concat, const int64_t *nt_rule_ptr,
                   const float *cum, int64_t start_id, int64_t max_length, int64_t max_tries, int64_t *out,
                   int64_t seed) {
    std::mt19937 gen(seed);
    std::uniform_real_distribution<float> dist(0.0f, 1.0f);
    auto unif = [&]() { return dist(gen); };
    return sample_row_with(rhs_ptr, rhs_concat, nt_rule_ptr, cum, start_id, max_length, max_tries, out, unif);
}

// ------------------------------------------------- infix -> postfix / prefix
// Length of the leading non-NO_OP run, or -1 if a token follows a NO_OP.
inline int32_t infix_length(const int64_t *row, int64_t M) {
    int32_t len = 0;
    bool ended = false;
    for (int64_t i = 0; i < M; ++i) {
        if (row[i] == kNoOp) ended = true;
        else if (ended) return -1;
        else ++len;
    }
    return len;
}

int64_t infix_to_postfix_row(const int64_t *prec, const int64_t *in, int64_t *out, int64_t lp, int64_t rp,
                             int64_t M) {
    const int32_t len = infix_length(in, M);
    if (len < 0) return E_PARSE_UNEXPECTED;
    int32_t ops[kMaxLen], q[kMaxLen];
    int32_t nops = 0, nq = 0;
    for (int32_t i = 0; i < len; ++i) {
        const int32_t tok = (int32_t)in[i];
        if (prec[tok] == 0) {
            if (nq >= kMaxLen) return E_PARSE_QUEUE;
            q[nq++] = tok;
        } else if (tok == lp) {
            if (nops >= kMaxLen) return E_PARSE_STACK;
            ops[nops++] = tok;
        } else if (tok == rp) {
            while (nops > 0 && ops[nops - 1] != lp) {
                if (nq >= kMaxLen) return E_PARSE_QUEUE;
                q[nq++] = ops[--nops];
            }
            if (nops == 0) return E_PARSE_PAREN;
            --nops;  // drop '('
        } else {
            while (nops > 0 && ops[nops - 1] != lp &&
                   ((left_assoc(tok) && prec[ops[nops - 1]] >= prec[tok]) ||
                    (right_assoc(tok) && prec[ops[nops - 1]] > prec[tok]))) {
                if (nq >= kMaxLen) return E_PARSE_QUEUE;
                q[nq++] = ops[--nops];
            }
            if (nops >= kMaxLen) return E_PARSE_STACK;
            ops[nops++] = tok;
        }
    }
    while (nops > 0) {
        if (ops[nops - 1] == lp) return E_PARSE_PAREN;
        if (nq >= kMaxLen) return E_PARSE_QUEUE;
        q[nq++] = ops[--nops];
    }
    if (nq > 0) {  // arity check of the postfix sequence
        int32_t depth = 0;
        for (int32_t i = 0; i < nq; ++i) {
            const int32_t tok = q[i];
            if (prec[tok] == 0) { ++depth; continue; }
            const int a = arity(tok);
            if (depth < a) return a == 1 ? E_TREE_UNARY : (a == 2 ? E_TREE_BINARY : E_INTERNAL);
            depth += 1 - a;
        }
        if (depth != 1) return E_TREE_MALFORMED;
    }
    for (int32_t i = 0; i < nq; ++i) out[i] = q[i];
    for (int64_t i = nq; i < M; ++i) out[i] = kNoOp;
    return E_OK;
}

int64_t infix_to_prefix_row(const int64_t *prec, const int64_t *in, int64_t *out, int64_t lp, int64_t rp,
                            int64_t M) {
    const int32_t len = infix_length(in, M);
    if (len < 0) return E_PARSE_UNEXPECTED;
    int32_t ops[kMaxLen], q[kMaxLen];
    int32_t nops = 0, nq = 0;
    for (int32_t i = len - 1; i >= 0; --i) {  // right to left, parentheses swap roles
        const int32_t tok = (int32_t)in[i];
        if (prec[tok] == 0) {
            if (nq >= kMaxLen) return E_PARSE_QUEUE;
            q[nq++] = tok;
        } else if (tok == rp) {
            if (nops >= kMaxLen) return E_PARSE_STACK;
            ops[nops++] = tok;
        } else if (tok == lp) {
            while (nops > 0 && ops[nops - 1] != rp) {
                if (nq >= kMaxLen) return E_PARSE_QUEUE;
                q[nq++] = ops[--nops];
            }
            if (nops == 0) return E_PARSE_PAREN;
            --nops;
        } else {
            while (nops > 0 && ops[nops - 1] != rp &&
                   ((left_assoc(tok) && prec[ops[nops - 1]] > prec[tok]) ||
                    (right_assoc(tok) && prec[ops[nops - 1]] >= prec[tok]))) {
                if (nq >= kMaxLen) return E_PARSE_QUEUE;
                q[nq++] = ops[--nops];
            }
            if (nops >= kMaxLen) return E_PARSE_STACK;
            ops[nops++] = tok;
        }
    }
    while (nops > 0) {
        if (ops[nops - 1] == rp) return E_PARSE_PAREN;
        if (nq >= kMaxLen) return E_PARSE_QUEUE;
        q[nq++] = ops[--nops];
    }
    for (int32_t i = 0; i < nq; ++i) out[i] = q[nq - 1 - i];
    for (int64_t i = nq; i < M; ++i) out[i] = kNoOp;
    return E_OK;
}

// ------------------------------------------------------ postfix -> infix
// The infix string is built right to left into a buffer of 3*kMaxLen tokens:
// reading the postfix backwards, an operator opens a frame (writes ')' and,
// for a postfix-style unary, the operator itself); a completed operand closes
// frames: a unary frame writes '(' (and its name if functional), a binary
// frame first writes its operator between the right and the left operand,
// then '(' once the left operand is complete.
int64_t postfix_to_infix_row(const int64_t *in, int64_t *out, int64_t lp, int64_t rp, int64_t M, int64_t Mi) {
    bool started = false;
    int32_t len = 0, balance = 0, roots = 0;
    for (int64_t k = M - 1; k >= 0; --k) {
        const int64_t tok = in[k];
        if (tok == kNoOp) {
            if (started) return E_CONV_UNEXPECTED;
            continue;
        }
        started = true;
        ++len;
        balance += arity(tok) - 1;
        if (balance == -1) ++roots;
    }
    if (len == 0) {
        for (int64_t k = 0; k < Mi; ++k) out[k] = kNoOp;
        return E_OK;
    }
    if (roots > 1) return E_CONV_ROOTS;
    enum { WANT_RIGHT = 0, WANT_LEFT = 1 };
    int32_t fop[kMaxLen], fstate[kMaxLen];
    int32_t nf = 0;
    int64_t buf[kMaxLen * 3];
    int32_t pos = kMaxLen * 3 - 1;  // next free slot, filled downwards
    for (int32_t idx = len - 1; idx >= 0; --idx) {
        const int64_t tok = in[idx];
        const int a = arity(tok);
        if (a == 0) {
            if (pos < 0) return E_CONV_STACK;
            buf[pos--] = tok;
            bool closed = true;
            while (closed && nf > 0) {
                const int32_t op = fop[nf - 1];
                if (arity(op) == 1) {
                    if (functional(op)) {
                        if (pos - 1 < 0) return E_CONV_STACK;
                        buf[pos--] = lp;
                        buf[pos--] = op;
                    } else {
                        if (pos < 0) return E_CONV_STACK;
                        buf[pos--] = lp;
                    }
                    --nf;
                } else if (fstate[nf - 1] == WANT_RIGHT) {
                    if (pos < 0) return E_CONV_STACK;
                    buf[pos--] = op;
                    fstate[nf - 1] = WANT_LEFT;
                    closed = false;
                } else {
                    if (pos < 0) return E_CONV_STACK;
                    buf[pos--] = lp;
                    --nf;
                }
            }
        } else {
            if (nf >= kMaxLen) return E_CONV_STACK;
            fop[nf] = (int32_t)tok;
            fstate[nf] = WANT_RIGHT;
            if (a == 1 && !functional(tok)) {
                if (pos < 0) return E_CONV_STACK;
                buf[pos--] = tok;
            }
            if (pos < 0) return E_CONV_STACK;
            buf[pos--] = rp;
            ++nf;
        }
    }
    if (nf > 0) return arity(fop[nf - 1]) == 1 ? E_CONV_UNARY : E_CONV_BINARY;
    const int32_t total = (kMaxLen * 3 - 1) - pos;
    if (total > Mi) return E_CONV_INFIX_LONG;
    for (int32_t k = 0; k < total; ++k) out[k] = buf[pos + 1 + k];
    for (int64_t k = total; k < Mi; ++k) out[k] = kNoOp;
    return E_OK;
}

// ------------------------------------------------------- prefix -> infix
int64_t prefix_to_infix_row(const int64_t *in, int64_t *out, int64_t lp, int64_t rp, int64_t M, int64_t Mi) {
    int32_t len = 0, balance = 0, roots = 0;
    bool ended = false;
    for (int64_t i = 0; i < M; ++i) {
        if (in[i] == kNoOp) { ended = true; continue; }
        if (ended) return E_CONV_UNEXPECTED;
        ++len;
        balance += arity(in[i]) - 1;
        if (balance == -1) ++roots;
    }
    if (len == 0) {
        for (int64_t k = 0; k < Mi; ++k) out[k] = kNoOp;
        return E_OK;
    }
    if (roots > 1) return E_CONV_ROOTS;
    enum { WANT_LEFT = 0, WANT_RIGHT = 1 };
    int32_t fop[kMaxLen], fstate[kMaxLen];
    int32_t nf = 0;
    int64_t w = 0;  // write position in out
    for (int32_t idx = 0; idx < len; ++idx) {
        const int64_t tok = in[idx];
        const int a = arity(tok);
        if (a == 0) {
            if (w >= Mi) return E_CONV_INFIX_LONG;
            out[w++] = tok;
            bool closed = true;
            while (closed && nf > 0) {
                const int32_t op = fop[nf - 1];
                if (arity(op) == 1) {
                    if (w >= Mi) return E_CONV_INFIX_LONG;
                    out[w++] = rp;
                    if (!functional(op)) {
                        if (w >= Mi) return E_CONV_INFIX_LONG;
                        out[w++] = op;
                    }
                    --nf;
                } else if (fstate[nf - 1] == WANT_LEFT) {
                    if (w >= Mi) return E_CONV_INFIX_LONG;
                    out[w++] = op;
                    fstate[nf - 1] = WANT_RIGHT;
                    closed = false;
                } else {
                    if (w >= Mi) return E_CONV_INFIX_LONG;
                    out[w++] = rp;
                    --nf;
                }
            }
        } else {
            if (nf >= kMaxLen) return E_CONV_STACK;
            fop[nf] = (int32_t)tok;
            fstate[nf] = WANT_LEFT;
            if (a == 1 && functional(tok)) {
                if (w + 2 > Mi) return E_CONV_INFIX_LONG;
                out[w++] = tok;
                out[w++] = lp;
            } else {
                if (w + 1 > Mi) return E_CONV_INFIX_LONG;
                out[w++] = lp;
            }
            ++nf;
        }
    }
    if (nf != 0) return arity(fop[nf - 1]) == 1 ? E_CONV_UNARY : E_CONV_BINARY;
    for (int64_t k = w; k < Mi; ++k) out[k] = kNoOp;
    return E_OK;
}

// ------------------------------------------- prefix <-> postfix reordering
// Walks the input in its reading order with a stack of (position, children
// still expected); a node is emitted once all its children are emitted.
int64_t reorder_row(const int64_t *in, int64_t *out, int64_t M, bool from_prefix) {
    int32_t pstack[kMaxLen], need[kMaxLen];
    int32_t np = 0, nout = 0, balance = 0, roots = 0;
    bool seen_end = false;
    std::vector<int64_t> tmp;
    if (!from_prefix) tmp.resize((size_t)M);
    for (int64_t step = 0; step < M; ++step) {
        const int64_t i = from_prefix ? step : M - 1 - step;
        const int64_t tok = in[i];
        if (tok == kNoOp) {
            if (from_prefix) seen_end = true;
            else if (seen_end) return E_CONV_UNEXPECTED;
            continue;
        }
        if (from_prefix) {
            if (seen_end) return E_CONV_UNEXPECTED;
        } else {
            seen_end = true;  // "started" for the backwards scan
        }
        const int a = arity(tok);
        balance += a - 1;
        if (balance == -1) ++roots;
        if (np >= kMaxLen) return E_CONV_STACK;
        pstack[np] = (int32_t)i;
        need[np++] = a;
        while (np > 0 && need[np - 1] == 0) {
            --np;
            if (nout >= M) return from_prefix ? E_CONV_POSTFIX_LONG : E_CONV_PREFIX_LONG;
            const int64_t v = in[pstack[np]];
            if (from_prefix) out[nout++] = v;
            else tmp[(size_t)nout++] = v;
            if (np > 0) --need[np - 1];
        }
    }
    if (np != 0) return arity(in[pstack[np - 1]]) == 1 ? E_CONV_UNARY : E_CONV_BINARY;
    if (roots != 1 && nout > 0) return E_CONV_ROOTS;
    if (from_prefix) {
        for (int64_t i = nout; i < M; ++i) out[i] = kNoOp;
    } else {
        for (int32_t i = 0; i < nout; ++i) out[i] = tmp[(size_t)(nout - 1 - i)];
        for (int64_t i = nout; i < M; ++i) out[i] = kNoOp;
    }
    return E_OK;
}

// --------------------------------------------------- parent / children
int64_t prefix_parent_row(const int64_t *in, int64_t *parent, int64_t M) {
    for (int64_t i = 0; i < M; ++i) parent[i] = kNull;
    int32_t st[kMaxLen];
    int32_t n = 0;
    bool started = false;
    for (int64_t i = M - 1; i >= 0; --i) {
        const int32_t tok = (int32_t)in[i];
        if (tok == kNoOp) {
            if (started) return E_PARSE_UNEXPECTED;
            continue;
        }
        started = true;
        const int a = arity(tok);
        if (a == 1) {
            if (n < 1) return E_TREE_UNARY;
            parent[st[--n]] = i;
        } else if (a == 2) {
            if (n < 2) return E_TREE_BINARY;
            parent[st[--n]] = i;
            parent[st[--n]] = i;
        }
        if (n >= kMaxLen) return E_TREE_STACK;
        st[n++] = (int32_t)i;
    }
    if (started && n != 1) return E_TREE_MALFORMED;
    return E_OK;
}

int64_t postfix_children_row(const int64_t *in, int64_t *children, int64_t M) {
    int32_t st[kMaxLen];
    int32_t n = 0, len = 0;
    bool ended = false;
    for (int64_t i = 0; i < M; ++i) {
        const int32_t tok = (int32_t)in[i];
        if (tok == kNoOp) { ended = true; continue; }
        ++len;
        if (ended) return E_PARSE_UNEXPECTED;
        const int a = arity(tok);
        if (a == 1) {
            if (n < 1) return E_TREE_UNARY;
            children[2 * i] = st[--n];
            children[2 * i + 1] = kNull;
        } else if (a == 2) {
            if (n < 2) return E_TREE_BINARY;
            const int32_t right = st[--n];
            const int32_t left = st[--n];
            children[2 * i] = left;
            children[2 * i + 1] = right;
        } else {
            children[2 * i] = kNull;
            children[2 * i + 1] = kNull;
        }
        if (n >= kMaxLen) return E_TREE_STACK;
        st[n++] = (int32_t)i;
    }
    if (len > 0 && n != 1) return E_TREE_MALFORMED;
    return E_OK;
}

}  // namespace

extern "C" {

// Orders grammar rules by left-hand-side id with std::sort over an index
// vector (the order of rules that share a left-hand side is whatever
// libstdc++'s introsort yields, as in the reference's grammar compiler,
// src/pcfg.cpp:362-372; it decides which rule a sampled number selects).
void sr_pcfg_sort_rules(const int64_t *lhs, int64_t n, int64_t *perm) {
    std::vector<int64_t> idx((size_t)n);
    std::iota(idx.begin(), idx.end(), 0);
    std::sort(idx.begin(), idx.end(), [lhs](int64_t a, int64_t b) { return lhs[a] < lhs[b]; });
    for (int64_t i = 0; i < n; ++i) perm[i] = idx[(size_t)i];
}

void sr_pcfg_sample(const int64_t *rule_lhs, const int64_t *rhs_ptr, const int64_t *rhs_concat,
                    const int64_t *nt_rule_ptr, const float *cum_probs, int64_t start_id, int64_t max_length,
                    int64_t max_tries, int64_t *out, int64_t *errors, const int64_t *seeds, int64_t B) {
    (void)rule_lhs;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t b = 0; b < B; ++b)
        errors[b] = sample_row(rhs_ptr, rhs_concat, nt_rule_ptr, cum_probs, start_id, max_length, max_tries,
                               out + b * max_length, seeds[b]);
}

// The device sampler's stream (Philox-4x32-10, row b = subsequence b of
// `seed`), on the host: the same rows sr_pcfg_sample_device draws.
void sr_pcfg_sample_philox(const int64_t *rhs_ptr, const int64_t *rhs_concat, const int64_t *nt_rule_ptr,
                           const float *cum_probs, int64_t start_id, int64_t max_length, int64_t max_tries,
                           int64_t *out, int64_t *errors, uint64_t seed, int64_t B) {
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t b = 0; b < B; ++b) {
        sr::PhiloxUniform unif(seed, (uint64_t)b);
        errors[b] = sample_row_with(rhs_ptr, rhs_concat, nt_rule_ptr, cum_probs, start_id, max_length, max_tries,
                                    out + b * max_length, unif);
    }
}

void sr_philox4x32_10(const uint32_t *ctr, const uint32_t *key, uint32_t *out) { sr::philox4x32_10(ctr, key, out); }

void sr_pcfg_infix_to_postfix(const int64_t *prec, const int64_t *in, int64_t *out, int64_t *errors, int64_t lp,
                              int64_t rp, int64_t B, int64_t M) {
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t b = 0; b < B; ++b) errors[b] = infix_to_postfix_row(prec, in + b * M, out + b * M, lp, rp, M);
}

void sr_pcfg_infix_to_prefix(const int64_t *prec, const int64_t *in, int64_t *out, int64_t *errors, int64_t lp,
                             int64_t rp, int64_t B, int64_t M) {
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t b = 0; b < B; ++b) errors[b] = infix_to_prefix_row(prec, in + b * M, out + b * M, lp, rp, M);
}

void sr_pcfg_postfix_to_infix(const int64_t *in, int64_t *out, int64_t *errors, int64_t lp, int64_t rp, int64_t B,
                              int64_t M, int64_t Mi) {
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t b = 0; b < B; ++b) errors[b] = postfix_to_infix_row(in + b * M, out + b * Mi, lp, rp, M, Mi);
}

void sr_pcfg_prefix_to_infix(const int64_t *in, int64_t *out, int64_t *errors, int64_t lp, int64_t rp, int64_t B,
                             int64_t M, int64_t Mi) {
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t b = 0; b < B; ++b) errors[b] = prefix_to_infix_row(in + b * M, out + b * Mi, lp, rp, M, Mi);
}

void sr_pcfg_prefix_to_postfix(const int64_t *in, int64_t *out, int64_t *errors, int64_t B, int64_t M) {
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t b = 0; b < B; ++b) errors[b] = reorder_row(in + b * M, out + b * M, M, true);
}

void sr_pcfg_postfix_to_prefix(const int64_t *in, int64_t *out, int64_t *errors, int64_t B, int64_t M) {
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t b = 0; b < B; ++b) errors[b] = reorder_row(in + b * M, out + b * M, M, false);
}

void sr_pcfg_get_prefix_parent(const int64_t *in, int64_t *parent, int64_t *errors, int64_t B, int64_t M) {
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t b = 0; b < B; ++b) errors[b] = prefix_parent_row(in + b * M, parent + b * M, M);
}

void sr_pcfg_get_postfix_children(const int64_t *in, int64_t *children, int64_t *errors, int64_t B, int64_t M) {
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t b = 0; b < B; ++b) errors[b] = postfix_children_row(in + b * M, children + b * M * 2, M);
}

}  // extern "C"
