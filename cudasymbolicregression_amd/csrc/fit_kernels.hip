// Device-resident batched L-BFGS for expression constants (the fit path).
//
// Algorithm and options follow the reference's LBFGS_batch.step with
// optimize's closure (symbolic_torch/lbfgs_batch.py:65-326,
// symbolic_torch/optimization.py:63-101), per problem:
//   * p = -H g by the two-loop recursion (rho = 1/(y.s + eps),
//     gamma = s.y/(y.y + eps) from the newest pair, no scaling before the
//     first pair);
//   * strong Wolfe bracketing (c1 = 1e-4, c2 = 0.9, alpha0 = 1,
//     alpha <- alpha + 0.8*(20 - alpha); a point meeting the curvature test is
//     accepted before the bracket tests) with the bisection zoom, or Armijo
//     backtracking (halving);
//   * x <- x + alpha p, history push (drop oldest beyond m);
//   * stop when ||g||_2 < tol_grad (NaN never stops).
// Closures the reference recomputes at points it already evaluated (phi_0 at
// x, phi_lo, and f at the accepted step) are reused: evaluation is
// deterministic, so the values are identical.  Inner products, the norm, the
// losses and the Wolfe tests run in FP64; vectors and step lengths stay FP32
// like the reference's tensors.  A problem stops early when it has converged,
// when its loss reaches the optional per-problem stop_loss, or when every
// slot of x is NaN with a non-empty history (the reference's iterates stay
// NaN from there on).
//
// Execution model (gfx950, wave64):
//   * A problem's constant slots are COMPACTED: lane j of a slot vector holds
//     the j-th LEARNABLE_CONSTANT node (and 64 + j when SL == 2).  Every
//     other node has gradient 0, so all of them follow one common trajectory
//     (their steps are +-0, or NaN once a non-finite scalar appears); lane
//     nc ("the shared slot") carries it, which keeps the reference's
//     NaN-propagation through the M-slot inner products and its all-NaN
//     stop test exact.
//   * The L-BFGS iteration is a state machine around ONE closure call site,
//     so the (large) evaluation code is instantiated once per kernel.
//   * Closure: one tile = 64 points, one per lane.  The tape (node values)
//     and the adjoints are [node][lane] arrays in LDS (no bank conflicts);
//     programs longer than 64 nodes keep them in a global workspace slot.
//     Node words are read with v_readlane, so the opcode dispatch is scalar.
//   * W > 1 waves may share one problem: points are split between them, the
//     per-wave FP64 partial sums meet in LDS in a fixed order, and every wave
//     runs the (deterministic) L-BFGS logic on identical values.
//   * Problems are bucketed by program length so every kernel is sized for
//     its bucket; the buckets run concurrently on forked streams.
#include "sr_code.h"
#include "sr_device_info.h"
#include "sr_queue.h"
#include "sr_streams.h"
#include "sr_eval.h"
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

using namespace sr;

namespace {

constexpr int MHMAX = 16;      // largest supported history_size
constexpr int FIT_NCREG = 8;   // constants accumulated in per-lane FP64 registers
constexpr int FIT_NB = 5;      // tape-row buckets

struct FitArgs {
    const float *X;
    const float *Y;
    int64_t N;
    int n_x;
    const int64_t *ops;
    const int64_t *ch;
    const float *C_init;
    int M;
    int64_t B;
    const int64_t *x_off;    // ragged: element offset of problem b's X rows (nullable)
    const int64_t *y_off;    // ragged: element offset of problem b's targets (nullable)
    const int32_t *n_pts;    // ragged: number of points of problem b (nullable)
    const double *stop_loss; // per-problem early stop threshold on the loss (nullable)
    float *C_out;
    double *loss_out;
    int32_t *iters_out;
    int32_t *nclos_out;      // closures evaluated per problem (nullable)
    int32_t *err_out;
    int32_t *flags;          // [0] validation (max), [1] evaluation (max)
    int max_iter;
    int history;
    double tol_grad;
    int line_search;         // 0 strong Wolfe, 1 backtracking
    int ls_max_iter;
    double eps;
    const int32_t *list;     // problems of this bucket
    const int32_t *list_count;
    int32_t *work_counter;
    float *gtape;            // global tape slots (programs longer than 64 nodes)
    int stationary_exit;     // end a problem once its iteration is an exact fixed point
    int starts;              // fits per problem: item b = problem * starts + start
    // Blocks past grid_cap (0: none) leave at once when the other buckets
    // hold at least 1/32 as many problems as this one (counts_all: the
    // problems per bucket)
    int grid_cap;
    const int32_t *counts_all;
};

// FP64 inner product over the K live slots; lanes past K never contribute.
// With K <= 4 only the first quad holds terms: its two quad steps and one
// + 0.0 give the bits of the full row reduction (whose other steps add +0).
template <int SL>
__device__ inline double sdot(const SVec<float, SL> &a, const SVec<float, SL> &b, int K) {
    const int lane = lane_id();
    double s = (lane < K) ? (double)a.v[0] * (double)b.v[0] : 0.0;
    if (SL > 1 && lane + WAVE < K) s += (double)a.v[SL - 1] * (double)b.v[SL - 1];
    if (K <= 4) {
        s += dpp_f64<DPP_QUAD_SWAP1>(s);
        s += dpp_f64<DPP_QUAD_SWAP2>(s);
        return uniform_f64(rl(s, 0) + 0.0);
    }
    return wave_sum_rows(s, K <= 16 ? 1 : 4);
}

// ------------------------------------------------------------ tile data
// Variable rows (v < NX, in tape rows LT + v) and targets of the tile group
// starting at point base.
template <int PPT, int NX, int LT>
__device__ inline void load_group(const float *X, const float *Yb, int64_t N, int n_x, int64_t base,
                                  const Tape<float, PPT> &t, float (&target)[PPT]) {
    const int lane = lane_id();
    float xv[PPT][NX];
#pragma unroll
    for (int p = 0; p < PPT; ++p) {
        const int64_t n = base + p * WAVE + lane;
        const bool act = n < N;
        const float *row = X + (act ? n : 0) * (int64_t)n_x;
#pragma unroll
        for (int v = 0; v < NX; ++v) xv[p][v] = (act && v < n_x) ? row[v] : 0.0f;
        target[p] = act ? Yb[n] : 0.0f;
    }
#pragma unroll
    for (int v = 0; v < NX; ++v)
        if (v < n_x)
#pragma unroll
            for (int p = 0; p < PPT; ++p) t.put(LT + v, p, xv[p][v]);
}

// ---------------------------------------------------------------- closure
// One wave's share of an objective evaluation at the compact constants xt:
// tile groups of 64*PPT points starting at group `wslot`, every W-th group.
// Returns the wave's FP64 loss sum (uniform), its FP64 dL/dC partials in
// compact lanes (gc), and its error code (uniform).  Per instruction there is
// one opcode dispatch for all PPT points; operands come from the plan
// (sr_code.h): a tape row, a scalar resolved once per evaluation, or a
// variable read from X (v >= NXR_ROWS); adjoints overwrite values in place.
//
// When the wave's points form one tile group (single), their variable rows
// were written once per problem (load_single_group) and their targets are
// in tg1; nothing is read from X or Y here.
template <int SL, int NX, int PPT, int LT>
__device__ inline void closure_code(const Code &c, const Plan &pl, int nc, const SVec<float, SL> &xt, const float *X,
                                    const float *Yb, int64_t N, int n_x, int wslot, int W, const Tape<float, PPT> &t,
                                    bool single, const float (&tg1)[PPT], double &loss, double (&gc)[SL],
                                    int32_t &err) {
    const int lane = lane_id();
    const int NI = c.NI;
    constexpr int HW = LT > WAVE ? 2 : 1;
    float sv0[HW], sv1[HW];
    resolve_scalars<float, SL, LT>(c, xt, sv0, sv1);
    auto scalar = [&](const float (&sv)[HW], int i) -> float {  // instruction i's scalar operand
        if constexpr (HW == 1) return rl(sv[0], i);
        else return i < WAVE ? rl(sv[0], i) : rl(sv[HW - 1], i - WAVE);
    };
    double lacc = 0.0;
    double cacc[FIT_NCREG];
#pragma unroll
    for (int q = 0; q < FIT_NCREG; ++q) cacc[q] = 0.0;
#pragma unroll
    for (int h = 0; h < SL; ++h) gc[h] = 0.0;
    int32_t e = 0;
    const float inv_n = 1.0f / (float)N;
    const int64_t group = (int64_t)WAVE * PPT;
    for (int64_t base = (int64_t)wslot * group; base < N; base += (int64_t)W * group) {
        bool act[PPT];
        int64_t n[PPT];
        float target[PPT];
#pragma unroll
        for (int p = 0; p < PPT; ++p) {
            n[p] = base + p * WAVE + lane;
            act[p] = n[p] < N;
            target[p] = tg1[p];
        }
        if (!single) load_group<PPT, NX, LT>(X, Yb, N, n_x, base, t, target);
        // zeroed extra adjoint rows
        for (int r = NI; r < NI + pl.nextra; ++r)
#pragma unroll
            for (int p = 0; p < PPT; ++p) t.put(r, p, 0.0f);
        // operands of instruction i: tape rows or scalars, both rows read
        // before either is used (a unary instruction reads row 0 and discards
        // it); variables v >= NX are read from X
        auto fetch = [&](uint32_t f, int i, float (&a0)[PPT], float (&a1)[PPT]) {
            const bool bin = pf_bin(f);
            const float s0 = scalar(sv0, i), s1 = scalar(sv1, i);
            if (pf_xg(f, 0) || pf_xg(f, 1)) {
                const uint32_t w = c.word_lt<LT>(i);
#pragma unroll
                for (int k = 0; k < 2; ++k) {
                    float *av = k == 0 ? a0 : a1;
                    const int v = desc_idx(k == 0 ? instr_d0(w) : instr_d1(w));
                    const bool xg = pf_xg(f, k), sc = pf_sc(f, k);
                    const int row = pf_row(f, k);
#pragma unroll
                    for (int p = 0; p < PPT; ++p) {
                        const float tv = xg ? (act[p] ? X[n[p] * (int64_t)n_x + v] : 0.0f) : t.get(row, p);
                        av[p] = (k == 1 && !bin) ? 0.0f : (sc ? (k == 0 ? s0 : s1) : tv);
                    }
                }
            } else {
                const int r0 = pf_row(f, 0), r1 = pf_row(f, 1);
                const bool c0 = pf_sc(f, 0), c1 = pf_sc(f, 1);
                float t0[PPT], t1[PPT];
#pragma unroll
                for (int p = 0; p < PPT; ++p) t0[p] = t.get(r0, p);
#pragma unroll
                for (int p = 0; p < PPT; ++p) t1[p] = t.get(r1, p);
#pragma unroll
                for (int p = 0; p < PPT; ++p) {
                    a0[p] = c0 ? s0 : t0[p];
                    a1[p] = bin ? (c1 ? s1 : t1[p]) : 0.0f;
                }
            }
        };
        // ---- forward
        for (int i = 0; i < NI; ++i) {
            const uint32_t f = pl.fw<LT>(i);
            const int op = pf_op(f);
            float a0[PPT], a1[PPT], r[PPT];
            fetch(f, i, a0, a1);
            if (op == OP_COPY) {
#pragma unroll
                for (int p = 0; p < PPT; ++p) r[p] = a0[p];
            } else {
                fwd_ops<float, PPT>(op, a0, a1, r);
            }
#pragma unroll
            for (int p = 0; p < PPT; ++p) t.put(i, p, r[p]);
        }
        // ---- loss and the MSE upstream gradient (2 * (y - Y)) * (1 / N),
        //      which replaces the root's value
#pragma unroll
        for (int p = 0; p < PPT; ++p) {
            const float d = t.get(NI - 1, p) - target[p];
            if (act[p]) lacc += (double)d * (double)d;
            t.put(NI - 1, p, act[p] ? (2.0f * d) * inv_n : 0.0f);
        }
        // ---- reverse sweep
        for (int i = NI - 1; i >= 0; --i) {
            const uint32_t f = pl.fw<LT>(i), sw = pl.sw<LT>(i);
            const int op = pf_op(f);
            const bool binary = pf_bin(f);
            float g[PPT], a0[PPT], a1[PPT], ge0[PPT], ge1[PPT];
            const int grow = ps_grow(sw);
#pragma unroll
            for (int p = 0; p < PPT; ++p) g[p] = t.get(grow, p);
            fetch(f, i, a0, a1);
            bwd_masked<float, PPT>(op, g, a0, a1, act, ge0, ge1, e);
            // scatter to the operands; ge is the amount added to an operand's
            // adjoint (0 where the reference adds nothing or past the end)
            auto scatter = [&](int k, const float (&ge)[PPT]) {
                const uint32_t mode = ps_mode(sw, k);
                const int tg = ps_tgt(sw, k);
                if (mode - 1u < 3u) {
                    // PM_WRITE: the only adjoint of a one-parent row, replacing
                    // its value; PM_CROW: a constant used once, summed after the
                    // tile group; PM_ACC: an extra adjoint row, accumulated
                    float o[PPT];
#pragma unroll
                    for (int p = 0; p < PPT; ++p) o[p] = 0.0f;
                    if (mode == PM_ACC)
#pragma unroll
                        for (int p = 0; p < PPT; ++p) o[p] = t.get(tg, p);
#pragma unroll
                    for (int p = 0; p < PPT; ++p) t.put(tg, p, o[p] + ge[p]);
                } else if (mode == PM_CONST) {  // a constant used several times: wave sum per use
                    double gv = 0.0;
#pragma unroll
                    for (int p = 0; p < PPT; ++p) gv += (double)ge[p];
                    const double s = wave_sum_rows(gv, 4);
                    if (tg < WAVE) { if (lane == tg) gc[0] += s; }
                    else if (lane == tg - WAVE) gc[SL - 1] += s;
                } else if (mode == PM_NOOP) {
#pragma unroll
                    for (int p = 0; p < PPT; ++p)
                        if (ge[p] != 0.0f) e = e > ERR_GRAD_NO_OP ? e : ERR_GRAD_NO_OP;
                }
            };
            scatter(0, ge0);
            if (binary) scatter(1, ge1);
        }
        // ---- adjoints of the constants used once, from their rows, in FP64
#pragma unroll
        for (int q = 0; q < FIT_NCREG; ++q) {
            if (q < nc) {
                const int row = pl.crow(q);
                if (row != (int)NO_CROW) {
                    double gv = 0.0;
#pragma unroll
                    for (int p = 0; p < PPT; ++p) gv += (double)t.get(row, p);
                    cacc[q] += gv;
                }
            }
        }
        for (int j = FIT_NCREG; j < nc; ++j) {
            const int row = pl.crow(j);
            if (row == (int)NO_CROW) continue;
            double gv = 0.0;
#pragma unroll
            for (int p = 0; p < PPT; ++p) gv += (double)t.get(row, p);
            const double s = wave_sum_rows(gv, 4);
            if (j < WAVE) { if (lane == j) gc[0] += s; }
            else if (lane == j - WAVE) gc[SL - 1] += s;
        }
    }
    loss = wave_sum_rows(lacc, 4);
    const int ncr = nc < FIT_NCREG ? nc : FIT_NCREG;
#pragma unroll
    for (int q = 0; q < FIT_NCREG; ++q) {
        if (q < ncr) {
            const double s = wave_sum_rows(cacc[q], 4);
            if (lane == q) gc[0] += s;
        }
    }
    err = wave_max_err(e);
}

// ----------------------------------------------------------------- kernel
// W == 1: a block holds block_waves independent waves, one problem each
// (4, or 2 with two slot registers per lane, whose history is twice as large).
// W > 1:  a block of W waves works on one problem.
__host__ __device__ constexpr int block_waves(int W, int SL) { return W == 1 ? (SL == 1 ? 4 : 2) : W; }

// Per-group LDS: history S, Y [hist][SL*64] floats, rho [MHMAX] doubles.
__host__ __device__ inline size_t hist_bytes(int hist, int SL) {
    return (size_t)2 * hist * SL * WAVE * sizeof(float) + (size_t)MHMAX * sizeof(double);
}
__host__ __device__ inline size_t part_bytes(int W, int SL) {
    return W > 1 ? (size_t)2 * W * (SL * WAVE + 2) * sizeof(double) : 0;
}
// Tape of one wave: LT instruction rows and NXR_ROWS variable rows per
// point of a tile group, 64 floats each.
__host__ __device__ inline size_t tape_bytes(int LT, int PPT, bool mem) {
    return mem ? 0 : tape_elems(LT, PPT) * sizeof(float);
}
__host__ __device__ inline size_t fit_lds(int LT, int PPT, bool mem, int W, int SL, int hist) {
    const int groups = W == 1 ? block_waves(W, SL) : 1;
    return (size_t)block_waves(W, SL) * tape_bytes(LT, PPT, mem) + (size_t)groups * hist_bytes(hist, SL) +
           part_bytes(W, SL) + 16;
}

// Internal error: a program reached a kernel whose tape it does not fit
// (the bucketing rule and the compiler disagree).  Never expected.
constexpr int32_t ERR_BUCKET = 900;

enum FitState : int { ST_INIT = 0, ST_SW = 1, ST_SW_ZERO = 2, ST_ZOOM = 3, ST_BT = 4, ST_FINAL = 5 };

// LT: most tape rows of a program in the bucket (count_rows_wave); NX:
// variables held in tape rows per point; SL: slot registers per lane; W: waves per problem;
// MEM: tape in the global workspace instead of LDS; WPE: waves per SIMD the
// register budget is sized for; PPT: points per lane per dispatch.
template <int LT, int NX, int SL, int W, bool MEM, int WPE, int PPT>
__global__ __launch_bounds__(block_waves(W, SL) * 64) __attribute__((amdgpu_waves_per_eu(WPE)))
void k_fit(FitArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int wave = uniform_i32((int)(threadIdx.x / WAVE));
    const int lane = lane_id();
    const int group = W == 1 ? wave : 0;   // problem slot of this wave inside the block
    const int wslot = W == 1 ? 0 : wave;   // this wave's share of the points
    const int hist = a.history;
    Tape<float, PPT> tape;
    {
        float *slot = MEM ? a.gtape + ((size_t)blockIdx.x * block_waves(W, SL) + wave) * tape_elems(LT, PPT)
                          : reinterpret_cast<float *>(smem) + (size_t)wave * tape_elems(LT, PPT);
        tape.v = slot;
    }
    unsigned char *hbase = smem + (size_t)block_waves(W, SL) * tape_bytes(LT, PPT, MEM);
    const int groups = W == 1 ? block_waves(W, SL) : 1;
    unsigned char *gbase = hbase + (size_t)group * hist_bytes(hist, SL);
    float *Sh = reinterpret_cast<float *>(gbase);
    float *Yh = Sh + (size_t)hist * SL * WAVE;
    double *rho_l = reinterpret_cast<double *>(Yh + (size_t)hist * SL * WAVE);
    double *part = reinterpret_cast<double *>(hbase + (size_t)groups * hist_bytes(hist, SL));
    int *next_slot = reinterpret_cast<int *>(hbase + (size_t)groups * hist_bytes(hist, SL) + part_bytes(W, SL));
    // every listed problem expands into `starts` consecutive work items, so
    // the fits of one problem run side by side over the same program,
    // targets and points
    if (a.grid_cap > 0 && (int)blockIdx.x >= a.grid_cap) {
        // The short-program bucket holds most problems; with its full
        // persistent grid its waves take the issue slots that the longer
        // buckets' slowest problems (which set the call's duration) need.
        int64_t other = 0;
        for (int k = 1; k < FIT_NB; ++k) other += a.counts_all[k];
        if (other * 32 >= (int64_t)a.counts_all[0]) return;  // uniform over the block
    }
    const int64_t count = (int64_t)*a.list_count * a.starts;
    const double c1 = 1e-4, c2 = 0.9;

    int buf = 0;

    for (;;) {
        int i;
        if (W == 1) {
            i = wave_fetch_add(a.work_counter);
        } else {
            if (threadIdx.x == 0) *next_slot = atomicAdd(a.work_counter, 1);
            __syncthreads();
            i = uniform_i32(*next_slot);
            __syncthreads();
        }
        if (i >= count) break;
        const int64_t pb = (int64_t)a.list[i / a.starts];       // problem
        const int64_t b = pb * a.starts + (int64_t)(i % a.starts);  // item: its constants and results
        const CProg<float> p = load_cprog<float>(a.ops, a.ch, a.C_init, pb, a.M, a.n_x, b);
        const bool writer = (wslot == 0);
        if (p.verr) {
            if (writer) {
                if (lane < a.M) a.C_out[b * a.M + lane] = p.cinit0;
                if (lane + WAVE < a.M) a.C_out[b * a.M + lane + WAVE] = p.cinit1;
                if (lane == 0) {
                    atomicMax(&a.flags[0], p.verr);
                    a.err_out[b] = p.verr;
                    a.iters_out[b] = 0;
                    if (a.nclos_out) a.nclos_out[b] = 0;
                    a.loss_out[b] = __builtin_nan("");
                }
            }
            continue;
        }
        const Code code = compile_code(p, a.M);
        const Plan plan = compile_plan(p, a.M, code, LT, NX);
        if (code.NI + plan.nextra > LT) {
            if (writer && lane == 0) {
                atomicMax(&a.flags[1], ERR_BUCKET);
                a.err_out[b] = ERR_BUCKET;
                a.iters_out[b] = 0;
                if (a.nclos_out) a.nclos_out[b] = 0;
                a.loss_out[b] = __builtin_nan("");
            }
            continue;
        }
        const float *X = a.X + (a.x_off ? a.x_off[pb] : 0);
        const float *Yb = a.Y + (a.y_off ? a.y_off[pb] : pb * a.N);
        const int64_t N = a.n_pts ? (int64_t)a.n_pts[pb] : a.N;
        const double stop = a.stop_loss ? a.stop_loss[pb] : -1.0;
        const int K = p.K;
        // the wave's points form one tile group: load them once per problem
        const bool single = N <= (int64_t)W * WAVE * PPT;
        float tg1[PPT];
#pragma unroll
        for (int q = 0; q < PPT; ++q) tg1[q] = 0.0f;
        if (single) load_group<PPT, NX, LT>(X, Yb, N, a.n_x, (int64_t)wslot * WAVE * PPT, tape, tg1);

        SVec<float, SL> x = compact_start<float, SL>(p), g, d, g_prev, g_lo, g_new, ga;
#pragma unroll
        for (int h = 0; h < SL; ++h) { g.v[h] = d.v[h] = g_prev.v[h] = g_lo.v[h] = g_new.v[h] = ga.v[h] = 0.0f; }
        int st = ST_INIT;
        float a_eval = 0.0f, ai = 1.0f, a_prev = 0.0f, lo = 0.0f, hi = 0.0f, alpha = 0.0f;
        double f = 0.0, f_new = 0.0, dphi0 = 0.0, phi_prev = 0.0, phi_lo = 0.0;
        int ls = 0, zi = 0, it = 0, nh = 0, head = 0, nclos = 0;
        int zrun = 0;          // newest history pairs with s == 0 and y == 0, consecutively
        bool fz_probe = false; // this line search starts from an all-zero direction
        bool d_finite = true;
        int32_t err = 0;

        for (;;) {
            // ---------------- one objective evaluation at x (+ a_eval d)
            SVec<float, SL> xt;
#pragma unroll
            for (int h = 0; h < SL; ++h) xt.v[h] = (st == ST_INIT) ? x.v[h] : x.v[h] + a_eval * d.v[h];
            double phi, gcd[SL];
            int32_t e2;
            closure_code<SL, NX, PPT, LT>(code, plan, p.nc, xt, X, Yb, N, a.n_x, wslot, W, tape, single, tg1, phi,
                                          gcd, e2);
            if (W > 1) {
                // partial sums of the W waves meet in LDS; every wave adds them
                // in the same order, so all waves hold identical results
                double *pb = part + (size_t)buf * W * (SL * WAVE + 2);
                double *mine = pb + (size_t)wslot * (SL * WAVE + 2);
                if (lane == 0) { mine[0] = phi; mine[1] = (double)e2; }
#pragma unroll
                for (int h = 0; h < SL; ++h) mine[2 + h * WAVE + lane] = gcd[h];
                __syncthreads();
                phi = 0.0;
                int32_t em = 0;
#pragma unroll
                for (int h = 0; h < SL; ++h) gcd[h] = 0.0;
                for (int w = 0; w < W; ++w) {
                    const double *o = pb + (size_t)w * (SL * WAVE + 2);
                    phi += o[0];
                    const int32_t ew = (int32_t)o[1];
                    em = ew > em ? ew : em;
#pragma unroll
                    for (int h = 0; h < SL; ++h) gcd[h] += o[2 + h * WAVE + lane];
                }
                phi = uniform_f64(phi);
                e2 = uniform_i32(em);
                buf ^= 1;
            }
            phi = phi / (double)N;
#pragma unroll
            for (int h = 0; h < SL; ++h) ga.v[h] = (float)gcd[h];
            ++nclos;
            if (e2) { err = e2; break; }
            if (fz_probe) {
                // All stored pairs are zero, so the direction is zero and x + d == x.
                // If the objective at x reproduces (f, g) bit for bit, this
                // iteration accepts alpha = 1 and leaves (x, f, g) unchanged while
                // pushing another zero pair: every remaining iteration repeats it.
                fz_probe = false;
                bool same = (phi == f);
#pragma unroll
                for (int h = 0; h < SL; ++h)
                    if (lane + h * WAVE < K && __float_as_uint(ga.v[h]) != __float_as_uint(g.v[h])) same = false;
                if (__ballot(!same) == 0ull) {
                    it = a.max_iter;
                    break;
                }
            }

            // ---------------- advance the iteration's state
            bool accept = false;  // (alpha, f_new, g_new) decided
            bool eval_final = false;
            bool zoom_next = false;
            if (st == ST_INIT) {
                f = phi;
                g = ga;
            } else if (st == ST_FINAL) {
                f_new = phi;
                g_new = ga;
                accept = true;
            } else if (st == ST_SW) {
                const double dphi = sdot<SL>(ga, d, K);
                const bool bad = (phi > f + c1 * (double)ai * dphi0) || (ls > 0 && phi >= phi_prev);
                const bool good = fabs(dphi) <= c2 * fabs(dphi0);
                if (good) {
                    alpha = ai; f_new = phi; g_new = ga; accept = true;
                } else if (bad || dphi >= 0.0) {
                    lo = bad ? a_prev : ai;
                    hi = bad ? ai : a_prev;
                    phi_lo = bad ? phi_prev : phi;
                    g_lo = bad ? g_prev : ga;
                    zi = 0;
                    if (bad && ls == 0 && !d_finite) {
                        st = ST_SW_ZERO;  // x + 0*d is not x: evaluate it
                        a_eval = 0.0f;
                        continue;
                    }
                    zoom_next = true;
                } else {
                    a_prev = ai;
                    phi_prev = phi;
                    g_prev = ga;
                    ai = ai + (20.0f - ai) * 0.8f;
                    ++ls;
                    if (ls < a.ls_max_iter) { a_eval = ai; continue; }
                    alpha = ai;
                    eval_final = true;
                }
            } else if (st == ST_SW_ZERO) {
                phi_lo = phi;
                g_lo = ga;
                zoom_next = true;
            } else if (st == ST_ZOOM) {
                const float aj = a_eval;
                const double dphij = sdot<SL>(ga, d, K);
                const bool zbad = (phi > f + c1 * (double)aj * dphi0) || (phi >= phi_lo);
                if (fabs(dphij) <= c2 * fabs(dphi0)) {
                    alpha = aj; f_new = phi; g_new = ga; accept = true;
                } else {
                    const float new_hi = zbad ? aj : hi;
                    const bool flip = dphij * (double)(new_hi - lo) >= 0.0;
                    hi = (!zbad && flip) ? lo : new_hi;
                    if (!zbad) { lo = aj; phi_lo = phi; g_lo = ga; }
                    if (fabsf(hi - lo) < 1e-9f) {
                        alpha = aj; f_new = phi; g_new = ga; accept = true;
                    } else {
                        ++zi;
                        zoom_next = true;
                    }
                }
            } else {  // ST_BT
                if (phi <= f + 1e-4 * (double)ai * dphi0) {
                    alpha = ai; f_new = phi; g_new = ga; accept = true;
                } else {
                    ai = ai * 0.5f;
                    ++ls;
                    if (ls < a.ls_max_iter) { a_eval = ai; continue; }
                    alpha = ai;
                    eval_final = true;
                }
            }
            if (zoom_next) {
                if (zi < 10) {
                    a_eval = (lo + hi) / 2.0f;
                    st = ST_ZOOM;
                    continue;
                }
                alpha = lo; f_new = phi_lo; g_new = g_lo; accept = true;
            }
            if (eval_final) {
                // no point was accepted: step by alpha and evaluate there
                st = ST_FINAL;
                a_eval = alpha;
                continue;
            }
            if (accept) {
                // ---------------- step and history push
                SVec<float, SL> s, yv;
#pragma unroll
                for (int h = 0; h < SL; ++h) {
                    s.v[h] = alpha * d.v[h];
                    x.v[h] = x.v[h] + s.v[h];
                    yv.v[h] = g_new.v[h] - g.v[h];
                }
                const double rho_new = 1.0 / (sdot<SL>(yv, s, K) + a.eps);
                int slot;
                if (nh == hist) {
                    slot = head;
                    head = (head + 1 == hist) ? 0 : head + 1;
                } else {
                    slot = head + nh < hist ? head + nh : head + nh - hist;
                    ++nh;
                }
#pragma unroll
                for (int h = 0; h < SL; ++h) {
                    Sh[((size_t)slot * SL + h) * WAVE + lane] = s.v[h];
                    Yh[((size_t)slot * SL + h) * WAVE + lane] = yv.v[h];
                }
                if (lane == 0) rho_l[slot] = rho_new;
                {
                    bool nz = false;
#pragma unroll
                    for (int h = 0; h < SL; ++h)
                        if (lane + h * WAVE < K && (s.v[h] != 0.0f || yv.v[h] != 0.0f)) nz = true;
                    zrun = __ballot(nz) == 0ull ? (zrun < hist ? zrun + 1 : hist) : 0;
                }
                g = g_new;
                f = f_new;
                ++it;
            }
            // ---------------- start of an iteration: tests, direction, line search
            if (it >= a.max_iter) break;
            const double gn = sqrt(sdot<SL>(g, g, K));
            if (gn < a.tol_grad) break;
            if (f <= stop) break;
            if (nh > 0) {
                bool notnan = false;
#pragma unroll
                for (int h = 0; h < SL; ++h) {
                    const int j = lane + h * WAVE;
                    if (j < K) {
                        const bool isn = (x.v[h] != x.v[h]) || (j == p.nc && p.shared_nan_init);
                        if (!isn) notnan = true;
                    }
                }
                if (__ballot(notnan) == 0ull) break;
            }
            // two-loop recursion, newest pair first; alpha_j kept in lane j.
            // Each pair's history rows are read one step ahead, so their LDS
            // latency overlaps the previous step's inner product.
            auto load_pair = [&](int jj, SVec<float, SL> &sj, SVec<float, SL> &yj, double &rj) {
                const int sl = head + jj < hist ? head + jj : head + jj - hist;
#pragma unroll
                for (int h = 0; h < SL; ++h) {
                    sj.v[h] = Sh[((size_t)sl * SL + h) * WAVE + lane];
                    yj.v[h] = Yh[((size_t)sl * SL + h) * WAVE + lane];
                }
                rj = rho_l[sl];
            };
            SVec<float, SL> q = g;
            double alph_lane = 0.0;
            SVec<float, SL> sn, yn;
            double rn = 0.0;
            if (nh > 0) load_pair(nh - 1, sn, yn, rn);
            for (int jj = nh - 1; jj >= 0; --jj) {
                const SVec<float, SL> sj = sn, yj = yn;
                const double rj = rn;
                if (jj > 0) load_pair(jj - 1, sn, yn, rn);
                const double aj = uniform_f64(rj) * sdot<SL>(sj, q, K);
                if (lane == jj) alph_lane = aj;
                const float af = (float)aj;
#pragma unroll
                for (int h = 0; h < SL; ++h) q.v[h] = q.v[h] - af * yj.v[h];
            }
            SVec<float, SL> z = q;
            if (nh > 0) {
                SVec<float, SL> s_new, y_new;  // the newest pair
                double r_new;
                load_pair(nh - 1, s_new, y_new, r_new);
                load_pair(0, sn, yn, rn);
                const float gamma = (float)(sdot<SL>(s_new, y_new, K) / (sdot<SL>(y_new, y_new, K) + a.eps));
#pragma unroll
                for (int h = 0; h < SL; ++h) z.v[h] = gamma * q.v[h];
                for (int jj = 0; jj < nh; ++jj) {
                    const SVec<float, SL> sj = sn, yj = yn;
                    const double rj = rn;
                    if (jj + 1 < nh) load_pair(jj + 1, sn, yn, rn);
                    const double beta = uniform_f64(rj) * sdot<SL>(yj, z, K);
                    const float coef = (float)(uniform_f64(rl(alph_lane, jj)) - beta);
#pragma unroll
                    for (int h = 0; h < SL; ++h) z.v[h] = z.v[h] + sj.v[h] * coef;
                }
            }
#pragma unroll
            for (int h = 0; h < SL; ++h) d.v[h] = -z.v[h];
            dphi0 = sdot<SL>(g, d, K);
            {
                bool bad = false;
#pragma unroll
                for (int h = 0; h < SL; ++h)
                    if (lane + h * WAVE < K && !isfinite(d.v[h])) bad = true;
                d_finite = __ballot(bad) == 0ull;
            }
            if (a.stationary_exit && nh > 0 && zrun >= nh) {
                // zero direction and no -0 in x (x + (+-0) == x bit for bit)
                bool moves = false;
#pragma unroll
                for (int h = 0; h < SL; ++h)
                    if (lane + h * WAVE < K && (d.v[h] != 0.0f || __float_as_uint(x.v[h]) == 0x80000000u))
                        moves = true;
                fz_probe = __ballot(moves) == 0ull;
            }
            ls = 0;
            ai = 1.0f;
            if (a.line_search == 0) {
                a_prev = 0.0f;
                phi_prev = f;
                g_prev = g;
                st = ST_SW;
            } else {
                st = ST_BT;
            }
            if (a.ls_max_iter <= 0) {
                alpha = ai;
                st = ST_FINAL;
            }
            a_eval = ai;
        }

        // ---------------- write back all M slots
        if (writer) {
            const float shared = slot_get<float, SL>(x, p.nc < 2 * WAVE ? p.nc : 0);
            const bool shared_nan = (K > p.nc) && (shared != shared);
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int k = lane + h * WAVE;
                const uint64_t cm = h == 0 ? p.cm0 : p.cm1;
                const uint64_t below = lane == 0 ? 0ull : (~0ull >> (WAVE - lane));
                const int o = (h == 0 ? 0 : __popcll(p.cm0)) + __popcll(cm & below);
                // every lane takes part in the shuffles
                const float v0 = __shfl(x.v[0], o & (WAVE - 1), WAVE);
                const float v1 = __shfl(x.v[SL - 1], o & (WAVE - 1), WAVE);
                if (k < a.M) {
                    float out;
                    if ((cm >> lane) & 1ull) out = (SL == 1 || o < WAVE) ? v0 : v1;
                    else out = shared_nan ? __builtin_nanf("") : (h == 0 ? p.cinit0 : p.cinit1);
                    a.C_out[b * a.M + k] = out;
                }
            }
            if (lane == 0) {
                a.loss_out[b] = f;
                a.iters_out[b] = it;
                if (a.nclos_out) a.nclos_out[b] = nclos;
                a.err_out[b] = err;
                if (err) atomicMax(&a.flags[1], err);
            }
        }
    }
}

// Buckets problems by tape rows (count_rows_wave: instructions plus extra
// adjoint rows) and live slot count K: rows <= 8, <= 16, <= 32, <= 64 (with
// K <= 64), and the rest.
constexpr int NB_CLASSIFY = 5;  // buckets of k_classify (== NBUCKETS below)
static_assert(NB_CLASSIFY == FIT_NB, "k_classify and the fit kernel disagree on the buckets");

// One wave per problem (the program row read coalesced); the run counts of
// the deterministic bucket lists follow in k_queue_count (sr_queue.h).
__global__ __launch_bounds__(256) void k_classify(const int64_t *__restrict__ ops, const int64_t *__restrict__ ch, int M,
                                                  int64_t B, int n_x, int8_t *bucket_of) {
    const int64_t b = (int64_t)blockIdx.x * 4 + uniform_i32((int)(threadIdx.x / WAVE));
    if (b >= B) return;
    int nc = 0;
    const int nr = count_rows_wave(ops + b * M, ch + b * M * 2, M, n_x, nc);
    const int K = nc + (M > nc ? 1 : 0);
    const int bucket = K > 64 ? 4 : (nr <= 8 ? 0 : (nr <= 16 ? 1 : (nr <= 32 ? 2 : (nr <= 64 ? 3 : 4))));
    if (lane_id() == 0) bucket_of[b] = (int8_t)bucket;
}

__global__ void k_fill_i32_f(int32_t *p, int32_t v, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ------------------------------------------------------------ launch plan
template <int LT, int NX, int SL, int W, bool MEM, int WPE, int PPT>
int launch_fit(const FitArgs &a, hipStream_t s, int grid) {
    const size_t lds = fit_lds(LT, PPT, MEM, W, SL, a.history);
    if (lds > lds_limit_device()) return -3;
    hipLaunchKernelGGL((k_fit<LT, NX, SL, W, MEM, WPE, PPT>), dim3((unsigned)grid), dim3(block_waves(W, SL) * 64),
                       lds, s, a);
    return (int)hipGetLastError();
}

// Persistent grid: as many blocks as the CUs can hold at once (cached per
// history size, which sets the LDS request).
template <int LT, int NX, int SL, int W, bool MEM, int WPE, int PPT>
int grid_for(int hist) {
    static int cache[MHMAX + 1] = {0};
    if (cache[hist] == 0) {
        int per_cu = 0;
        const size_t lds = fit_lds(LT, PPT, MEM, W, SL, hist);
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)k_fit<LT, NX, SL, W, MEM, WPE, PPT>,
                                                         block_waves(W, SL) * 64, lds) != hipSuccess ||
            per_cu <= 0)
            per_cu = 1;
        cache[hist] = num_cus() * per_cu;
    }
    return cache[hist];
}

// Bucket kernels <LT, NX, SL, W, MEM, WPE, PPT>.  Each bucket has a few
// launch variants (waves per problem W x points per lane PPT); each call
// names the one it runs per bucket (sr_fit_options, from the autotuner) or
// takes the built-in default.
constexpr int NXR = NXR_ROWS;  // variables in tape rows (the rest are read from X)
constexpr int NBUCKETS = 5;
static_assert(NBUCKETS == NB_CLASSIFY, "k_classify and the launch plan disagree on the buckets");
constexpr int B4_BLOCKS_PER_CU = 2;  // x 2 waves: 1024 long problems in flight

inline int b4_grid() { return num_cus() * B4_BLOCKS_PER_CU; }

template <int LT, int SL, int W, bool MEM, int WPE, int PPT>
int launch_variant(const FitArgs &a, hipStream_t s) {
    const int grid = MEM ? b4_grid() : grid_for<LT, NXR, SL, W, MEM, WPE, PPT>(a.history);
    return launch_fit<LT, NXR, SL, W, MEM, WPE, PPT>(a, s, grid);
}

struct FitVariant {
    int (*launch)(const FitArgs &, hipStream_t);
    int waves;  // waves per problem
    int ppt;    // points per lane per dispatch
};

const FitVariant kB0[] = {{launch_variant<8, 1, 1, false, 3, 2>, 1, 2},
                          {launch_variant<8, 1, 2, false, 3, 2>, 2, 2},
                          {launch_variant<8, 1, 4, false, 3, 2>, 4, 2},
                          {launch_variant<8, 1, 1, false, 2, 4>, 1, 4},
                          {launch_variant<8, 1, 8, false, 2, 1>, 8, 1}};
const FitVariant kB1[] = {{launch_variant<16, 1, 1, false, 3, 2>, 1, 2},
                          {launch_variant<16, 1, 2, false, 3, 2>, 2, 2},
                          {launch_variant<16, 1, 4, false, 3, 2>, 4, 2},
                          {launch_variant<16, 1, 8, false, 2, 1>, 8, 1}};
const FitVariant kB2[] = {{launch_variant<32, 1, 2, false, 3, 1>, 2, 1},
                          {launch_variant<32, 1, 4, false, 3, 1>, 4, 1},
                          {launch_variant<32, 1, 4, false, 3, 2>, 4, 2},
                          {launch_variant<32, 1, 8, false, 2, 1>, 8, 1}};
const FitVariant kB3[] = {{launch_variant<64, 1, 4, false, 2, 1>, 4, 1},
                          {launch_variant<64, 1, 2, false, 2, 1>, 2, 1},
                          {launch_variant<64, 1, 4, false, 2, 2>, 4, 2},
                          {launch_variant<64, 1, 8, false, 2, 1>, 8, 1}};
const FitVariant kB4[] = {{launch_variant<256, 2, 1, true, 2, 1>, 1, 1}};

const FitVariant *const kVariants[NBUCKETS] = {kB0, kB1, kB2, kB3, kB4};
constexpr int kNumVariants[NBUCKETS] = {5, 4, 4, 4, 1};
// Fastest on the benchmark workloads (autotune.tune_fit); constant.  Bucket 2
// takes eight waves of one point per lane: equal on 16384 problems of 512
// shared points, 11% faster on 32768 problems with their own 64..4096 points.
constexpr int kDefaultVariant[NBUCKETS] = {2, 2, 3, 0, 0};

// The selected variant, or the bucket's first one when the selected one
// needs more LDS than a workgroup may have at this history size.
inline int launch_bucket(int k, const int *variant, const FitArgs &a, hipStream_t s) {
    const int r = kVariants[k][variant[k]].launch(a, s);
    return r == -3 ? kVariants[k][0].launch(a, s) : r;
}

// Workspace: lists [5][B] int32 | counters [5] counts + [5] work | queue
// classification (sr_queue.h) | global tape (bucket 4).
inline size_t fit_q_off(int64_t B) { return (((size_t)NBUCKETS * B * 4 + 64) + 255) / 256 * 256; }
inline size_t fit_gt_off(int64_t B) { return fit_q_off(B) + (queue_ws_bytes(B, NBUCKETS) + 255) / 256 * 256; }
inline int64_t fit_ws(int64_t B, int64_t M) {
    size_t off = fit_gt_off(B);
    if (M > 32)  // more than 64 tape rows (at most 2 M - 1) or live slots
        off += (size_t)b4_grid() * block_waves(1, 2) * tape_elems(256, 1) * sizeof(float);
    return (int64_t)off + 256;
}

}  // namespace

extern "C" {

int64_t sr_lbfgs_workspace_bytes(int64_t B, int64_t M) { return fit_ws(B, M); }

// Launch-variant control of the fit kernels (see include/sr_eval.h).
int sr_fit_num_buckets(void) { return NBUCKETS; }
int sr_fit_num_variants(int bucket) { return (bucket < 0 || bucket >= NBUCKETS) ? -1 : kNumVariants[bucket]; }
int sr_fit_default_variant(int bucket) {
    return (bucket < 0 || bucket >= NBUCKETS) ? -1 : kDefaultVariant[bucket];
}
int sr_fit_variant_info(int bucket, int variant, int *waves, int *ppt) {
    if (bucket < 0 || bucket >= NBUCKETS || variant < 0 || variant >= kNumVariants[bucket]) return -1;
    if (waves) *waves = kVariants[bucket][variant].waves;
    if (ppt) *ppt = kVariants[bucket][variant].ppt;
    return 0;
}

}  // extern "C"

namespace {

// One fit launch over B problems with `starts` fits each (see
// sr_lbfgs_fit_multi in include/sr_eval.h).
int fit_impl(const float *X, const float *Y, int64_t N, int64_t n_x, const int64_t *ops, const int64_t *ch,
             const float *C_init, int64_t M, int64_t B, int64_t starts, const int64_t *x_off, const int64_t *y_off,
             const int32_t *n_pts, const double *stop_loss, float *C_out, double *loss_out, int32_t *iters_out,
             int32_t *nclos_out, int32_t *err_out, int32_t *flags, int max_iter, int history, double tol_grad,
             int line_search, int ls_max_iter, double eps, const sr_fit_options *opts, void *workspace,
             int64_t ws_bytes, void *stream) {
    hipStream_t s = (hipStream_t)stream;
    if (M <= 0 || M > 128 || B < 0 || n_x < 0 || n_x > MAX_NX_CODE + 1 || history < 1 || history > MHMAX ||
        max_iter < 0 || ls_max_iter < 0 || (line_search != 0 && line_search != 1))
        return -1;
    if (starts < 1 || B * starts > 0x7fffffff) return -1;
    int variant[NBUCKETS];
    for (int k = 0; k < NBUCKETS; ++k) {
        const int v = opts ? opts->variant[k] : -1;
        if (v >= kNumVariants[k]) return -1;
        variant[k] = v < 0 ? kDefaultVariant[k] : v;
    }
    const int stationary_exit = opts ? (opts->stationary_exit != 0) : 1;
    if (n_pts == nullptr && N <= 0) return -1;
    if (ws_bytes < sr_lbfgs_workspace_bytes(B, M)) return -2;
    hipLaunchKernelGGL(k_fill_i32_f, dim3(1), dim3(64), 0, s, flags, 0, (int64_t)2);
    if (B == 0) return (int)hipGetLastError();
    char *ws = (char *)workspace;
    int32_t *lists = (int32_t *)ws;
    int32_t *counters = (int32_t *)(ws + (size_t)NBUCKETS * B * 4);  // [0..4] counts, [5..9] work
    hipLaunchKernelGGL(k_fill_i32_f, dim3(1), dim3(64), 0, s, counters, 0, (int64_t)(2 * NBUCKETS));
    void *qws = ws + fit_q_off(B);
    hipLaunchKernelGGL(k_classify, dim3((unsigned)((B + 3) / 4)), dim3(256), 0, s, ops, ch, (int)M, B, (int)n_x,
                       queue_bucket_of(qws));
    queue_finish(qws, B, NBUCKETS, lists, counters, s);
    FitArgs a;
    memset(&a, 0, sizeof(a));
    a.X = X; a.Y = Y; a.N = N; a.n_x = (int)n_x; a.ops = ops; a.ch = ch; a.C_init = C_init; a.M = (int)M; a.B = B;
    a.x_off = x_off; a.y_off = y_off; a.n_pts = n_pts; a.stop_loss = stop_loss;
    a.C_out = C_out; a.loss_out = loss_out; a.iters_out = iters_out; a.nclos_out = nclos_out; a.err_out = err_out;
    a.flags = flags;
    a.max_iter = max_iter; a.history = history; a.tol_grad = tol_grad; a.line_search = line_search;
    a.ls_max_iter = ls_max_iter; a.eps = eps;
    a.stationary_exit = stationary_exit;
    a.starts = (int)starts;
    a.gtape = M > 32 ? (float *)(ws + fit_gt_off(B)) : nullptr;
    auto bucket_args = [&](int k) {
        FitArgs r = a;
        r.list = lists + (size_t)k * B;
        r.list_count = counters + k;
        r.work_counter = counters + NBUCKETS + k;
        // the short-program bucket runs one block per CU beside longer ones
        r.grid_cap = k == 0 ? num_cus() : 0;
        r.counts_all = counters;
        return r;
    };
    int r = (int)hipGetLastError();
    if (r) return r;
    // A program has at most 2 M - 1 tape rows and M live slots, so short
    // rows only need the first buckets.
    if (M <= 4) return launch_bucket(0, variant, bucket_args(0), s);
    // Buckets 4+3, 2 and 1 on side streams, bucket 0 on the caller's stream,
    // joined back before returning (also after a failed launch).
    ForkScope fk(s);
    if ((r = fk.fork())) return fk.finish(r);
    // Buckets 4 and 3 share a stream.  With M <= 64, bucket 4 only takes
    // programs with extra adjoint rows (shared or unused nodes, rare), so it
    // goes second: a persistent grid that is not yet resident holds back the
    // next kernel on its stream.
    if (M > 64 && (r = launch_bucket(4, variant, bucket_args(4), fk.side(0)))) return fk.finish(r);
    if (M > 16 && (r = launch_bucket(3, variant, bucket_args(3), fk.side(0)))) return fk.finish(r);
    if (M > 32 && M <= 64 && (r = launch_bucket(4, variant, bucket_args(4), fk.side(0)))) return fk.finish(r);
    if (M > 8 && (r = launch_bucket(2, variant, bucket_args(2), fk.side(1)))) return fk.finish(r);
    if ((r = launch_bucket(1, variant, bucket_args(1), fk.side(2)))) return fk.finish(r);
    if ((r = launch_bucket(0, variant, bucket_args(0), s))) return fk.finish(r);
    if ((r = fk.finish(0))) return r;
    return (int)hipGetLastError();
}

}  // namespace

extern "C" {

// Fits the LEARNABLE_CONSTANT slots of B batch-first expressions to their
// targets.  See include/sr_eval.h for the argument contract.
int sr_lbfgs_fit(const float *X, const float *Y, int64_t N, int64_t n_x, const int64_t *ops, const int64_t *ch,
                 const float *C_init, int64_t M, int64_t B, const int64_t *x_off, const int64_t *y_off,
                 const int32_t *n_pts, const double *stop_loss, float *C_out, double *loss_out, int32_t *iters_out,
                 int32_t *nclos_out, int32_t *err_out, int32_t *flags, int max_iter, int history, double tol_grad,
                 int line_search, int ls_max_iter, double eps, const sr_fit_options *opts, void *workspace,
                 int64_t ws_bytes, void *stream) {
    return fit_impl(X, Y, N, n_x, ops, ch, C_init, M, B, 1, x_off, y_off, n_pts, stop_loss, C_out, loss_out,
                    iters_out, nclos_out, err_out, flags, max_iter, history, tol_grad, line_search, ls_max_iter, eps,
                    opts, workspace, ws_bytes, stream);
}

// K fits per problem in one launch.  See include/sr_eval.h.
int sr_lbfgs_fit_multi(const float *X, const float *Y, int64_t N, int64_t n_x, const int64_t *ops,
                       const int64_t *ch, const float *C_init, int64_t M, int64_t B, int64_t starts,
                       const int64_t *x_off, const int64_t *y_off, const int32_t *n_pts, const double *stop_loss,
                       float *C_out, double *loss_out, int32_t *iters_out, int32_t *nclos_out, int32_t *err_out,
                       int32_t *flags, int max_iter, int history, double tol_grad, int line_search, int ls_max_iter,
                       double eps, const sr_fit_options *opts, void *workspace, int64_t ws_bytes, void *stream) {
    return fit_impl(X, Y, N, n_x, ops, ch, C_init, M, B, starts, x_off, y_off, n_pts, stop_loss, C_out, loss_out,
                    iters_out, nclos_out, err_out, flags, max_iter, history, tol_grad, line_search, ls_max_iter, eps,
                    opts, workspace, ws_bytes, stream);
}

}  // extern "C"
