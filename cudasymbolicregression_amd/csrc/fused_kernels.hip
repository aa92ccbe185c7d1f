// Fused root-only forward + backward over batch-first layouts.
//
// What optimize's closure needs (symbolic_torch/optimization.py:63-101) and
// what the reference gets from evaluate + gather + backward, without the
// (M,N,B) cache: per expression b and point n the root value
// (root = #non-NO_OP - 1, optimization.py:55-58), and the reverse sweep from
// the root with upstream gradient
//   MODE 0: dy[b][n]                  (an upstream dL/dy),
//   MODE 1: (2 * (y - Y)) * (1 / N)   (the MSE closure; dy is the target Y,
//                                      loss_out[b] = mean_n (y - Y)^2 in FP64).
// Outputs: y (B,N), grad_C (B,M) summed over the points in FP64, grad_X
// (N,n_x) summed over the batch, per-expression error codes.  Ragged form:
// expression b reads its own n_pts[b] points at X + x_off[b] and its dy / y
// at offset y_off[b] (no grad_X).
//
// Execution (gfx950): one wave per expression; the program is compiled once
// per wave into an instruction list and an evaluation plan (sr_code.h) and
// evaluated over tile groups of 64*PPT points with the tape in LDS (values,
// and adjoints in place).  Expressions are bucketed by tape rows (<= 8,
// <= 16, <= 32, <= 64, the rest) so each bucket's kernel is sized for its
// tape; the buckets run concurrently on forked streams.  Expressions come
// from a shared work counter, except when grad_X is wanted: then every wave
// takes a fixed, cost-balanced segment of its bucket list (sr_queue.h) and
// sums its variable adjoints into a wave-private partial row (in LDS, or in
// the workspace when the rows do not fit), with plain adds in a fixed order.
// A block adds its waves' LDS rows in wave order into one partial row; two
// final kernels add all partial rows in FP64 in row order.  No float
// atomics: grad_X is the same bit for bit on every call.
#include "sr_code.h"
#include "sr_device_info.h"
#include "sr_queue.h"
#include "sr_streams.h"
#include "sr_eval.h"
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <mutex>

using namespace sr;

namespace {

constexpr int F_NCREG = 8;  // constants accumulated in per-lane FP64 registers
constexpr int F_NX = 4;     // variables held in tape rows per point (NXR_ROWS)
static_assert(F_NX == NXR_ROWS, "the variables of a point are the tape's variable rows");

// With want_gx, the adjoints of the variable rows accumulate in NXR_ROWS more
// tape rows, LT + NXR_ROWS + v, when their numbers fit the 8-bit scatter
// target (every bucket but the longest); the tape grows by those rows.
__host__ __device__ constexpr bool f_var_rows(int LT) { return LT + 2 * NXR_ROWS <= 256; }
__host__ __device__ inline size_t f_tape_elems(int LT, int PPT) {
    return (size_t)(LT + (f_var_rows(LT) ? 2 : 1) * NXR_ROWS) * PPT * WAVE;
}
constexpr int F_NBUCKETS = 5;
// Wave-private grad_X rows of a block are kept in LDS only when they are
// small: larger rows would cut the waves per CU (LDS rows of 12 KB per wave
// made the 65536 x 1024 x 3 evaluation 27% slower than rows in the
// workspace, which stay mostly in the caches).
constexpr size_t GX_LDS_ROWS_MAX = 8 * 1024;
constexpr size_t GX_GLOBAL_BUDGET = 256u << 20;  // wave-private rows in the workspace, at most
// Expressions per work-queue ticket: short programs are cheap enough that one
// atomic per expression would serialise the queue; longer ones take one each
// so that the last expressions spread over all waves.
__host__ __device__ constexpr int fetch_chunk(int LT) { return LT <= 8 ? 4 : (LT <= 16 ? 2 : 1); }

struct FArgs {
    const void *X;
    int64_t N;
    int n_x;
    const int64_t *ops;
    const int64_t *ch;
    const void *C;
    int M;
    int64_t B;
    const void *dy;
    void *y_out;        // (B,N) or nullptr
    double *loss_out;   // (B) MODE 1, or nullptr
    void *grad_C;       // (B,M) unless fwd_only
    void *gx_partial;   // [rows][N*n_x] partial rows (grad_X wanted), or nullptr
    int64_t gx_row0;    // first partial row of this launch
    int gx_lds;         // 1: wave rows in LDS, one partial row per block; 0: one per wave
    int64_t gx_rows;    // k_fused_wide: partial rows of this bucket (list chunks)
    const int32_t *seg; // static segments of this bucket's list (grad_X wanted), else nullptr
    int32_t *err_b;     // (B) or nullptr
    int32_t *flags;     // [0] validation (max), [1] backward (max)
    const int32_t *list;
    const int32_t *list_count;
    int32_t *work_counter;
    void *gtape;        // global tape slots (longest bucket)
    int mode;
    int fwd_only;
    int want_gx;
    const int64_t *x_off;   // ragged: element offset of expression b's X rows (nullable)
    const int64_t *y_off;   // ragged: offset of expression b's dy / y values
    const int32_t *n_pts;   // ragged: number of points of expression b
    // wide evaluator (k_fused_wide) only
    const uint32_t *plan;   // compiled programs, PREC_WORDS words each (k_plan_fused)
    const float *xt;        // X transposed: variable v's N values at xt + v * n_pad, zero-padded
    int64_t n_pad;          // N rounded up to a multiple of WIDE_PAD
    int vec;                // dy and y may be accessed 16 bytes at a time (N % 4 == 0, aligned)
    double *gx_tot;         // grad_X: per-program running FP64 sums, WIDE_TOTW per program
};

// LT: most instructions of an expression in the bucket; SL: slot registers
// per lane; PPT: points per lane per dispatch; WPB: waves per block; MEM:
// tape in the global workspace; GX: more variables than F_NX registers.
template <typename T, int LT, int SL, int PPT, int WPB, bool MEM, bool GX>
__global__ __launch_bounds__(WPB * 64) void k_fused_code(FArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int wave = uniform_i32((int)(threadIdx.x / WAVE));
    const int lane = lane_id();
    const T *X = (const T *)a.X;
    const T *dy = (const T *)a.dy;
    T *y_out = (T *)a.y_out;
    T *grad_C = (T *)a.grad_C;
    const int64_t N = a.N;
    const int n_x = a.n_x;
    const int64_t gx_len = N * n_x;
    const bool want_gx = a.want_gx != 0 && !a.fwd_only;
    const bool gx_lds = want_gx && a.gx_lds;
    const size_t gx_bytes = gx_lds ? (((size_t)WPB * gx_len * sizeof(T) + 15) / 16 * 16) : 0;
    Tape<T, PPT> t;
    {
        T *slot = MEM ? (T *)a.gtape + ((size_t)blockIdx.x * WPB + wave) * f_tape_elems(LT, PPT)
                      : reinterpret_cast<T *>(smem + gx_bytes) + (size_t)wave * f_tape_elems(LT, PPT);
        t.v = slot;
    }
    const int64_t count = (int64_t)*a.list_count;
    const T *const X0 = X;
    // work: a fixed segment of the list (grad_X wanted), else chunks of CH
    // expressions from the shared counter (one atomic per chunk)
    const int64_t gw = (int64_t)blockIdx.x * WPB + wave;
    int32_t seg_i = a.seg ? uniform_i32(a.seg[gw]) : 0;
    const int32_t seg_e = a.seg ? uniform_i32(a.seg[gw + 1]) : 0;
    // this wave's grad_X partial row, variable-major (element (n, v) at
    // v * N + n, so a wave's 64 points are 64 consecutive words); each element
    // is only ever touched by the lane that owns its point (n = lane mod 64),
    // zeroing included.  A row in the workspace whose segment is empty is
    // left alone: the row sums skip it.
    T *gxrow = nullptr;
    if (want_gx && (gx_lds || seg_i < seg_e)) {
        gxrow = gx_lds ? reinterpret_cast<T *>(smem) + (size_t)wave * gx_len
                      : (T *)a.gx_partial + (a.gx_row0 + (int64_t)blockIdx.x * WPB + wave) * gx_len;
        for (int v = 0; v < n_x; ++v)
            for (int64_t nz = lane; nz < N; nz += WAVE) gxrow[(int64_t)v * N + nz] = T(0);
    }
    constexpr int CH = fetch_chunk(LT);
    int32_t chunk = 0, left = 0;
    for (;;) {
        int i;
        if (a.seg) {
            if (seg_i >= seg_e) break;
            i = seg_i++;
        } else {
            if (left == 0) {
                chunk = wave_fetch_add_n(a.work_counter, CH);
                left = CH;
            }
            i = chunk + (CH - left);
            --left;
            if (i >= count) break;
        }
        const int64_t b = (int64_t)a.list[i];
        const CProg<T> p = load_cprog<T>(a.ops, a.ch, (const T *)a.C, b, a.M, n_x);
        if (p.verr) {
            if (lane == 0) {
                atomicMax(&a.flags[0], p.verr);
                if (a.err_b) a.err_b[b] = p.verr;
            }
            continue;
        }
        const Code c = compile_code(p, a.M);
        constexpr int NVR = GX ? 0 : NXR_ROWS;  // variables held in tape rows LT + v
        // their adjoints in rows LT + NXR_ROWS + v (VROWS), or in gxr registers
        constexpr bool VROWS = !GX && f_var_rows(LT);
        const int var_adj_row0 = (VROWS && want_gx) ? LT + NXR_ROWS : -1;
        // the variables the program reads (wave-uniform): only their grad_X
        // rows are touched
        int vmask = 0;
        if (!GX && want_gx) {
#pragma unroll
            for (int v = 0; v < NXR_ROWS; ++v) {
                const int vop = VAR_START_ID + v;
                const bool u0 = lane < a.M && node_op(p.w0) == vop;
                const bool u1 = lane + WAVE < a.M && node_op(p.w1) == vop;
                vmask |= (v < n_x && (__ballot(u0) | __ballot(u1))) ? 1 << v : 0;
            }
        }
        const Plan pl = compile_plan(p, a.M, c, LT, NVR, var_adj_row0);
        const int NI = c.NI;
        if (NI + pl.nextra > LT) {  // the bucketing rule and the compiler disagree: never expected
            if (lane == 0) {
                atomicMax(&a.flags[1], 900);
                if (a.err_b) a.err_b[b] = 900;
            }
            continue;
        }
        const SVec<T, SL> xc = compact_start<T, SL>(p);
        constexpr int HW = LT > WAVE ? 2 : 1;
        T sv0[HW], sv1[HW];
        resolve_scalars<T, SL, LT>(c, xc, sv0, sv1);
        auto scalar = [&](const T (&sv)[HW], int ii) -> T {  // instruction ii's scalar operand
            if constexpr (HW == 1) return rl(sv[0], ii);
            else return ii < WAVE ? rl(sv[0], ii) : rl(sv[HW - 1], ii - WAVE);
        };
        // this expression's points (the shared X, or its own rows when ragged)
        const int64_t Nb = a.n_pts ? (int64_t)a.n_pts[b] : N;
        const int64_t yo = a.y_off ? a.y_off[b] : b * N;
        X = X0 + (a.x_off ? a.x_off[b] : 0);
        const T inv_n = T(1.0) / T(Nb);
        double lacc = 0.0;
        double cacc[F_NCREG];
#pragma unroll
        for (int q = 0; q < F_NCREG; ++q) cacc[q] = 0.0;
        double gc[SL];
#pragma unroll
        for (int h = 0; h < SL; ++h) gc[h] = 0.0;
        int32_t e = 0;
        // The variables and upstream values of tile group base + 64 * PPT are
        // loaded while group base is evaluated, so no group waits on its loads.
        // Every load is issued unconditionally at an in-bounds address (points
        // past the end read the last point; without an upstream array the
        // loads read X) and masked afterwards: a fixed number of loads per
        // group lets the first use wait for the current group's loads only.
        const bool need_dy = a.mode == 1 || !a.fwd_only;
        const T *const dsrc = need_dy ? dy + yo : X;
        T xv[PPT][F_NX], dv[PPT], xn[PPT][F_NX], dn[PPT];
        auto load_group = [&](int64_t gb, T (&xo)[PPT][F_NX], T (&dout)[PPT]) {
#pragma unroll
            for (int q = 0; q < PPT; ++q) {
                const int64_t nq = gb + q * WAVE + lane;
                const bool aq = nq < Nb;
                const int64_t nl = aq ? nq : Nb - 1;
                const T *row = X + nl * (int64_t)n_x;
#pragma unroll
                for (int v = 0; v < F_NX; ++v) {
                    const bool use = !GX && v < n_x;
                    const T xval = row[use ? v : 0];
                    xo[q][v] = (use && aq) ? xval : T(0);
                }
                const T dval = dsrc[need_dy ? nl : 0];
                dout[q] = (need_dy && aq) ? dval : T(0);
            }
        };
        auto take_next = [&]() {
#pragma unroll
            for (int q = 0; q < PPT; ++q) {
                dv[q] = dn[q];
#pragma unroll
                for (int v = 0; v < F_NX; ++v) xv[q][v] = xn[q][v];
            }
        };
        if (Nb > 0) load_group(0, xv, dv);
        for (int64_t base = 0; base < Nb; base += (int64_t)WAVE * PPT, take_next()) {
            bool act[PPT];
            int64_t n[PPT];
            T gxr[PPT][VROWS ? 1 : F_NX];
#pragma unroll
            for (int q = 0; q < PPT; ++q) {
                n[q] = base + q * WAVE + lane;
                act[q] = n[q] < Nb;
#pragma unroll
                for (int v = 0; v < (VROWS ? 1 : F_NX); ++v) gxr[q][v] = T(0);
            }
            // this group's grad_X partial sums, read now and added to after
            // the reverse sweep, so the read's latency hides behind the
            // group's evaluation (the row is this wave's alone)
            T gxp[PPT][F_NX];
            if (!GX && want_gx) {
#pragma unroll
                for (int q = 0; q < PPT; ++q)
#pragma unroll
                    for (int v = 0; v < F_NX; ++v)
                        gxp[q][v] = (((vmask >> v) & 1) && act[q]) ? gxrow[(int64_t)v * N + n[q]] : T(0);
            }
            load_group(base + (int64_t)WAVE * PPT, xn, dn);
            // variable rows, and zeroed extra adjoint rows
#pragma unroll
            for (int v = 0; v < NVR; ++v)
                if (v < n_x)
#pragma unroll
                    for (int q = 0; q < PPT; ++q) t.put(LT + v, q, xv[q][v]);
            for (int r = NI; r < NI + pl.nextra; ++r)
#pragma unroll
                for (int q = 0; q < PPT; ++q) t.put(r, q, T(0));
            if (var_adj_row0 >= 0)
#pragma unroll
                for (int v = 0; v < NVR; ++v)
                    if ((vmask >> v) & 1)
#pragma unroll
                        for (int q = 0; q < PPT; ++q) t.put(var_adj_row0 + v, q, T(0));
            // operands of instruction ii: tape rows or scalars, both rows read
            // before either is used (a unary instruction reads row 0 and
            // discards it); with GX, variables read from X
            auto fetch = [&](uint32_t f, int ii, T (&a0)[PPT], T (&a1)[PPT]) {
                const bool bin = pf_bin(f);
                if (GX && (pf_xg(f, 0) || pf_xg(f, 1))) {
                    const uint32_t w = c.word_lt<LT>(ii);
#pragma unroll
                    for (int k = 0; k < 2; ++k) {
                        T *av = k == 0 ? a0 : a1;
                        const T sv = k == 0 ? scalar(sv0, ii) : scalar(sv1, ii);
                        const int v = desc_idx(k == 0 ? instr_d0(w) : instr_d1(w));
                        const bool xg = pf_xg(f, k), sc = pf_sc(f, k);
                        const int row = pf_row(f, k);
#pragma unroll
                        for (int q = 0; q < PPT; ++q) {
                            const T tv = xg ? (act[q] ? X[n[q] * (int64_t)n_x + v] : T(0)) : t.get(row, q);
                            av[q] = (k == 1 && !bin) ? T(0) : (sc ? sv : tv);
                        }
                    }
                } else {
                    const int r0 = pf_row(f, 0), r1 = pf_row(f, 1);
                    const bool c0 = pf_sc(f, 0), c1 = pf_sc(f, 1);
                    const T s0 = scalar(sv0, ii), s1 = scalar(sv1, ii);
                    T t0[PPT], t1[PPT];
#pragma unroll
                    for (int q = 0; q < PPT; ++q) t0[q] = t.get(r0, q);
#pragma unroll
                    for (int q = 0; q < PPT; ++q) t1[q] = t.get(r1, q);
#pragma unroll
                    for (int q = 0; q < PPT; ++q) {
                        a0[q] = c0 ? s0 : t0[q];
                        a1[q] = bin ? (c1 ? s1 : t1[q]) : T(0);
                    }
                }
            };
            // ---- forward
            for (int ii = 0; ii < NI; ++ii) {
                const uint32_t f = pl.fw<LT>(ii);
                const int op = pf_op(f);
                T a0[PPT], a1[PPT], r[PPT];
                fetch(f, ii, a0, a1);
                if (op == OP_COPY) {
#pragma unroll
                    for (int q = 0; q < PPT; ++q) r[q] = a0[q];
                } else {
                    fwd_ops<T, PPT>(op, a0, a1, r);
                }
#pragma unroll
                for (int q = 0; q < PPT; ++q) t.put(ii, q, r[q]);
            }
            // ---- root value, loss and upstream gradient (dy is read only in
            //      MODE 1 or when the backward runs; it may be null otherwise);
            //      the root's adjoint replaces its value
#pragma unroll
            for (int q = 0; q < PPT; ++q) {
                const T yh = t.get(NI - 1, q);
                if (act[q] && y_out) y_out[yo + n[q]] = yh;
                T g0 = T(0);
                if (a.mode == 1) {
                    const T d = yh - dv[q];
                    if (act[q]) lacc += (double)d * (double)d;
                    g0 = (T(2.0) * d) * inv_n;
                } else if (!a.fwd_only) {
                    g0 = dv[q];
                }
                t.put(NI - 1, q, act[q] ? g0 : T(0));
            }
            if (a.fwd_only) continue;
            // ---- reverse sweep
            for (int ii = NI - 1; ii >= 0; --ii) {
                const uint32_t f = pl.fw<LT>(ii), sw = pl.sw<LT>(ii);
                const int op = pf_op(f);
                const bool binary = pf_bin(f);
                T g[PPT], a0[PPT], a1[PPT], ge0[PPT], ge1[PPT];
                const int grow = ps_grow(sw);
#pragma unroll
                for (int q = 0; q < PPT; ++q) g[q] = t.get(grow, q);
                fetch(f, ii, a0, a1);
                bwd_masked<T, PPT>(op, g, a0, a1, act, ge0, ge1, e);
                // ge: the amount added to the operand's adjoint (0 where the
                // reference adds nothing or the point is past the end)
                auto scatter = [&](int k, const T (&ge)[PPT]) {
                    const uint32_t mode = ps_mode(sw, k);
                    const int tg = ps_tgt(sw, k);
                    if (mode - 1u < 3u) {
                        // PM_WRITE: the only adjoint of a one-parent row,
                        // replacing its value; PM_CROW: a constant used once,
                        // summed after the tile group; PM_ACC: an extra adjoint
                        // row, accumulated
                        T o[PPT];
#pragma unroll
                        for (int q = 0; q < PPT; ++q) o[q] = T(0);
                        if (mode == PM_ACC)
#pragma unroll
                            for (int q = 0; q < PPT; ++q) o[q] = t.get(tg, q);
#pragma unroll
                        for (int q = 0; q < PPT; ++q) t.put(tg, q, o[q] + ge[q]);
                    } else if (mode == PM_CONST) {  // a constant used several times: wave sum per use
                        double gv = 0.0;
#pragma unroll
                        for (int q = 0; q < PPT; ++q) gv += (double)ge[q];
                        const double s = wave_sum_rows(gv, 4);
                        if (tg < WAVE) { if (lane == tg) gc[0] += s; }
                        else if (lane == tg - WAVE) gc[SL - 1] += s;
                    } else if (mode == PM_VAR) {
                        if (want_gx) {
                            if constexpr (!GX) {
                                if constexpr (!VROWS) {
#pragma unroll
                                    for (int q = 0; q < PPT; ++q)
#pragma unroll
                                        for (int v = 0; v < F_NX; ++v)
                                            if (v == tg) gxr[q][v] += ge[q];
                                }
                            } else {
                                const uint32_t w = c.word_lt<LT>(ii);
                                const int v = desc_idx(k == 0 ? instr_d0(w) : instr_d1(w));
#pragma unroll
                                for (int q = 0; q < PPT; ++q)
                                    if (ge[q] != T(0)) gxrow[(int64_t)v * N + n[q]] += ge[q];
                            }
                        }
                    } else if (mode == PM_NOOP) {
#pragma unroll
                        for (int q = 0; q < PPT; ++q)
                            if (ge[q] != T(0)) e = e > ERR_GRAD_NO_OP ? e : ERR_GRAD_NO_OP;
                    }
                };
                scatter(0, ge0);
                if (binary) scatter(1, ge1);
            }
            // ---- adjoints of the constants used once, from their rows, in FP64
#pragma unroll
            for (int r = 0; r < F_NCREG; ++r) {
                if (r < p.nc) {
                    const int row = pl.crow(r);
                    if (row != (int)NO_CROW) {
                        double gv = 0.0;
#pragma unroll
                        for (int q = 0; q < PPT; ++q) gv += (double)t.get(row, q);
                        cacc[r] += gv;
                    }
                }
            }
            for (int j = F_NCREG; j < p.nc; ++j) {
                const int row = pl.crow(j);
                if (row == (int)NO_CROW) continue;
                double gv = 0.0;
#pragma unroll
                for (int q = 0; q < PPT; ++q) gv += (double)t.get(row, q);
                const double s = wave_sum_rows(gv, 4);
                if (j < WAVE) { if (lane == j) gc[0] += s; }
                else if (lane == j - WAVE) gc[SL - 1] += s;
            }
            // ---- variable adjoints of this tile group
            if (!GX && want_gx) {
#pragma unroll
                for (int q = 0; q < PPT; ++q) {
                    if (!act[q]) continue;
#pragma unroll
                    for (int v = 0; v < F_NX; ++v) {
                        if (!((vmask >> v) & 1)) continue;
                        T gxv = T(0);
                        if constexpr (VROWS) {
                            gxv = t.get(LT + NXR_ROWS + v, q);
                        } else {
                            gxv = gxr[q][v];
                        }
                        // a zero contribution is not added (the sum keeps a -0)
                        gxrow[(int64_t)v * N + n[q]] = gxv == T(0) ? gxp[q][v] : gxp[q][v] + gxv;
                    }
                }
            }
        }
        if (a.mode == 1 && a.loss_out) {
            const double s = wave_sum_rows(lacc, 4);
            if (lane == 0) a.loss_out[b] = s / (double)Nb;
        }
        if (a.fwd_only) {
            if (lane == 0 && a.err_b) a.err_b[b] = 0;
            continue;
        }
        const int ncr = p.nc < F_NCREG ? p.nc : F_NCREG;
#pragma unroll
        for (int r = 0; r < F_NCREG; ++r) {
            if (r < ncr) {
                const double s = wave_sum_rows(cacc[r], 4);
                if (lane == r) gc[0] += s;
            }
        }
        // dL/dC back to node slots: constant node k takes its compact lane
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int k = lane + h * WAVE;
            const uint64_t cm = h == 0 ? p.cm0 : p.cm1;
            const uint64_t below = lane == 0 ? 0ull : (~0ull >> (WAVE - lane));
            const int o = (h == 0 ? 0 : __popcll(p.cm0)) + __popcll(cm & below);
            const double v0 = __shfl(gc[0], o & (WAVE - 1), WAVE);
            const double v1 = __shfl(gc[SL - 1], o & (WAVE - 1), WAVE);
            if (k < a.M) grad_C[b * a.M + k] = ((cm >> lane) & 1ull) ? (T)((SL == 1 || o < WAVE) ? v0 : v1) : T(0);
        }
        e = wave_max_err(e);
        if (lane == 0) {
            if (e) atomicMax(&a.flags[1], e);
            if (a.err_b) a.err_b[b] = e;
        }
    }
    if (gx_lds) {  // the block's partial row: its waves' rows added in wave order
        __syncthreads();
        const T *rows = reinterpret_cast<const T *>(smem);
        T *dst = (T *)a.gx_partial + (a.gx_row0 + blockIdx.x) * gx_len;
        for (int64_t i = threadIdx.x; i < gx_len; i += WPB * WAVE) {
            T v = rows[i];
            for (int w = 1; w < WPB; ++w) v += rows[(size_t)w * gx_len + i];
            dst[i] = v;
        }
    }
}

// ---------------------------------------------------------------------------
// Planned programs and the wide evaluator (FP32, shared X of at most
// NXR_ROWS variables, programs of at most PLAN_MAX_ROWS tape rows).
//
// k_plan_fused compiles every short program once, one wave per program at
// full occupancy, into a record that the bucket kernels load with two
// instructions instead of compiling it again (the compilation is a chain of
// dependent wave-uniform steps, slow at the bucket kernels' occupancy).  It
// also validates and classifies every program, so Ops and Ch are read once.
//
// Record (PREC_WORDS 32-bit words): 0-15 the fetch words of instructions
// 0..15, 16-31 their scatter words, 32-47 and 48-63 their two scalar
// operands (float bits), 64-67 the tape row of constants 0..15 when used
// once (one byte each, NO_CROW otherwise), 68 = NI | nextra << 8 | nc << 16
// | used-variable mask << 24, 69-70 and 71-72 the constant-node masks of
// nodes 0-63 and 64-127.  Tape rows: instructions and extra rows in
// [0, LT), variable v in row LT + v (the block's copy of X, or the wave's
// rows); variable adjoints go to registers (PM_VAR).
//
// k_fused_wide evaluates a tile group of 64 x PPT points per dispatch with
// each lane's PPT points held as PPT / 4 float4 quads (four consecutive
// points), so a tape row is read and written 16 bytes per lane.
// ---------------------------------------------------------------------------
constexpr int PREC_WORDS = 80;
constexpr int PLAN_MAX_ROWS = 16;
constexpr int PLAN_MAX_NC = 16;
constexpr int WIDE_PAD = 1024;  // n_pad granularity: a multiple of every tile group (64 x PPT, PPT <= 16)
constexpr int64_t WIDE_CHUNKS_PER_WAVE = 2;  // grad_X list chunks (and partial rows) per wave

// One wave per program.  Programs that fail validation get their code
// (flags[0], err_b) and no bucket (-1).  With `plan`, programs of at most
// PLAN_MAX_ROWS rows and PLAN_MAX_NC constants are compiled into their
// record and go to bucket 0 (<= 8 rows, LT 8) or 1 (LT 16); the others
// follow the rule of k_classify_fused over buckets 2-4.
__global__ __launch_bounds__(256) void k_plan_fused(const int64_t *__restrict__ ops, const int64_t *__restrict__ ch,
                                                    const float *__restrict__ C, int M, int64_t B, int n_x,
                                                    int8_t *bucket_of, int32_t *cost_of, uint32_t *plan,
                                                    int32_t *err_b, int32_t *flags) {
    const int64_t b = (int64_t)blockIdx.x * 4 + uniform_i32((int)(threadIdx.x / WAVE));
    if (b >= B) return;
    const int lane = lane_id();
    const CProg<float> p = load_cprog<float>(ops, ch, C, b, M, n_x);
    if (p.verr) {
        if (lane == 0) {
            bucket_of[b] = (int8_t)-1;
            if (cost_of) cost_of[b] = 1;
            atomicMax(&flags[0], p.verr);
            if (err_b) err_b[b] = p.verr;
        }
        return;
    }
    const Code c = compile_code(p, M);
    const int rows = c.NI + extra_rows(p, M);
    const bool planned = rows <= PLAN_MAX_ROWS && p.nc <= PLAN_MAX_NC;
    const int bucket = planned ? (rows <= 8 ? 0 : 1) : (p.nc > 64 ? 4 : (rows <= 32 ? 2 : (rows <= 64 ? 3 : 4)));
    if (lane == 0) {
        bucket_of[b] = (int8_t)bucket;
        if (cost_of) cost_of[b] = rows + 2;
    }
    if (!planned) return;
    const int LT = bucket == 0 ? 8 : 16;
    const Plan pl = compile_plan(p, M, c, LT, NXR_ROWS, -1);
    const SVec<float, 1> xc = compact_start<float, 1>(p);
    float sv0[1], sv1[1];
    resolve_scalars<float, 1, PLAN_MAX_ROWS>(c, xc, sv0, sv1);
    int vmask = 0;
#pragma unroll
    for (int v = 0; v < NXR_ROWS; ++v) {
        const int vop = VAR_START_ID + v;
        const bool u0 = lane < M && node_op(p.w0) == vop;
        const bool u1 = lane + WAVE < M && node_op(p.w1) == vop;
        vmask |= (v < n_x && (__ballot(u0) | __ballot(u1))) ? 1 << v : 0;
    }
    uint32_t *rec = plan + (size_t)b * PREC_WORDS;
    const int src = lane & 15, part = lane >> 4;
    const uint32_t fw = (uint32_t)__shfl((int)pl.f0, src, WAVE);
    const uint32_t sw = (uint32_t)__shfl((int)pl.s0, src, WAVE);
    const uint32_t s0 = __float_as_uint(__shfl(sv0[0], src, WAVE));
    const uint32_t s1 = __float_as_uint(__shfl(sv1[0], src, WAVE));
    rec[lane] = part == 0 ? fw : (part == 1 ? sw : (part == 2 ? s0 : s1));
    // constant j's row is byte j of words 64..67
    const uint32_t cr = pl.crow0 & 255u;
    uint32_t cw = 0u;
#pragma unroll
    for (int k = 0; k < 4; ++k) cw |= ((uint32_t)__shfl((int)cr, (4 * lane + k) & (WAVE - 1), WAVE) & 255u) << (8 * k);
    const uint32_t h0 = (uint32_t)c.NI | ((uint32_t)pl.nextra << 8) | ((uint32_t)p.nc << 16) | ((uint32_t)vmask << 24);
    uint32_t hv = cw;
    hv = lane == 4 ? h0 : hv;
    hv = lane == 5 ? (uint32_t)p.cm0 : hv;
    hv = lane == 6 ? (uint32_t)(p.cm0 >> 32) : hv;
    hv = lane == 7 ? (uint32_t)p.cm1 : hv;
    hv = lane == 8 ? (uint32_t)(p.cm1 >> 32) : hv;
    if (lane < 9) rec[64 + lane] = hv;
}

// X (N, n_x) -> xt [NXR_ROWS][n_pad], variable-major, zero past N and n_x.
__global__ void k_xt_f32(const float *__restrict__ X, int64_t N, int n_x, int64_t n_pad, float *__restrict__ xt) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= NXR_ROWS * n_pad) return;
    const int64_t v = i / n_pad, n = i - v * n_pad;
    xt[i] = (v < n_x && n < N) ? X[n * n_x + v] : 0.0f;
}

__device__ inline float4 f4_zero() { return make_float4(0.f, 0.f, 0.f, 0.f); }
__device__ inline float f4_get(const float4 &v, int e) { return e == 0 ? v.x : (e == 1 ? v.y : (e == 2 ? v.z : v.w)); }

// Four values of a row of n floats starting at point n0 (zero past n): one
// 16-byte load when the caller allows it and the quad is whole.
__device__ inline float4 load_quad(const float *__restrict__ row, int64_t n0, int64_t n, bool vec) {
    if (vec && n0 + 3 < n) return *reinterpret_cast<const float4 *>(row + n0);
    float4 r;
    r.x = n0 < n ? row[n0] : 0.f;
    r.y = n0 + 1 < n ? row[n0 + 1] : 0.f;
    r.z = n0 + 2 < n ? row[n0 + 2] : 0.f;
    r.w = n0 + 3 < n ? row[n0 + 3] : 0.f;
    return r;
}
__device__ inline void store_quad(float *__restrict__ row, int64_t n0, int64_t n, bool vec, float4 v) {
    if (vec && n0 + 3 < n) {
        *reinterpret_cast<float4 *>(row + n0) = v;
        return;
    }
    if (n0 < n) row[n0] = v.x;
    if (n0 + 1 < n) row[n0 + 1] = v.y;
    if (n0 + 2 < n) row[n0 + 2] = v.z;
    if (n0 + 3 < n) row[n0 + 3] = v.w;
}

// LDS of the wide evaluator.  X for all points, variable-major ([v][n_pad],
// one copy per block shared by its waves) when it takes at most
// WIDE_XBLK_MAX bytes; otherwise each wave keeps its tile group's variables
// in tape rows LT + v.  Then each wave's tape: its rows of 64 x PPT points.
constexpr int64_t WIDE_XBLK_MAX = 32 * 1024;
__host__ __device__ inline bool wide_xblk(int n_x, int64_t n_pad) {
    return (int64_t)n_x * n_pad * (int64_t)sizeof(float) <= WIDE_XBLK_MAX;
}
__host__ __device__ inline int wide_rows(int LT, int n_x, int64_t n_pad) { return LT + (wide_xblk(n_x, n_pad) ? 0 : n_x); }
__host__ __device__ inline size_t wide_xblk_bytes(int n_x, int64_t n_pad) {
    return wide_xblk(n_x, n_pad) ? (size_t)n_x * n_pad * sizeof(float) : 0;
}

// Running FP64 sums of a program across the tile groups of a grad_X call
// (k_fused_wide, group-major): its nc constants, then its loss, then its
// error code, in a workspace of WIDE_TOTW doubles per program (nc <= 16).
constexpr int WIDE_TOTW = 18;

// LT: tape rows of the bucket (8 or 16); PPT: points per lane per dispatch
// (a multiple of 4); WPB: waves per block.  Same per-point arithmetic and
// reference rules as k_fused_code (fwd_ops, bwd_masked, the scatter modes),
// so y and the error codes are the same bit for bit; the FP64 sums (loss,
// grad_C) and the grad_X partial rows add the same terms in another order.
// Variable adjoints of a tile group accumulate in registers (PM_VAR).
//
// Without grad_X a wave evaluates one program over all its tile groups at a
// time.  With grad_X the list is cut into gx_rows chunks of consecutive
// positions; a wave takes a chunk from the shared counter and runs it group
// by group: every program of the chunk on tile group 0, then on group 1, and
// so on, so the chunk's variable adjoints of a group are summed in registers
// (program by program, in list order) and written once to the chunk's own
// partial row, while each program carries its FP64 sums from group to group
// in the workspace (a program's sums do not depend on its chunk).  Which
// wave runs a chunk does not change its row, so grad_X is the same on every
// call while the chunks balance the load.
template <int LT, int PPT, int WPB>
__global__ __launch_bounds__(WPB * 64) void k_fused_wide(FArgs a) {
    static_assert(PPT % 4 == 0, "points per lane come in quads");
    constexpr int Q = PPT / 4;
    constexpr int NCW = 4;  // constants summed in per-lane FP64 registers (the others per tile group)
    extern __shared__ __align__(16) unsigned char smem[];
    const int wave = uniform_i32((int)(threadIdx.x / WAVE));
    const int lane = lane_id();
    const int n_x = a.n_x;
    const int64_t N = a.N, NP = a.n_pad;
    const bool want_gx = a.want_gx != 0 && !a.fwd_only;
    const bool vec = a.vec != 0;
    const bool xblk = wide_xblk(n_x, NP);
    const int rows = wide_rows(LT, n_x, NP);
    float4 *const xb4 = reinterpret_cast<float4 *>(smem);
    float4 *const tape = reinterpret_cast<float4 *>(smem + wide_xblk_bytes(n_x, NP)) + (size_t)wave * rows * Q * WAVE;
    const float *const dy = (const float *)a.dy;
    float *const y_out = (float *)a.y_out;
    float *const grad_C = (float *)a.grad_C;
    const int64_t gx_len = N * n_x;
    const int64_t NP4 = NP / 4;
    const int64_t GRP = (int64_t)WAVE * PPT;
    if (xblk) {  // the block's copy of X
        const float4 *src = reinterpret_cast<const float4 *>(a.xt);
        for (int64_t z = threadIdx.x; z < n_x * NP4; z += WPB * WAVE) xb4[z] = src[(z / NP4) * NP4 + z % NP4];
        __syncthreads();
    }
    // this wave's variables of one tile group in rows LT + v (no block copy)
    const bool one_group = N <= GRP;
    auto put_var_rows = [&](int64_t base, int mask) {
        for (int v = 0; v < n_x; ++v) {
            if (!((mask >> v) & 1)) continue;
            const float4 *src = reinterpret_cast<const float4 *>(a.xt + (int64_t)v * NP + base);
#pragma unroll
            for (int q = 0; q < Q; ++q) tape[((LT + v) * Q + q) * WAVE + lane] = src[q * WAVE + lane];
        }
    };
    if (!xblk && one_group) put_var_rows(0, (1 << n_x) - 1);
    const bool need_dy = a.mode == 1 || !a.fwd_only;
    const float inv_n = 1.0f / (float)N;

    // ---- one program on one tile group.  In: its plan words (rw, hw).  Out
    // (added to): lg per-lane FP64 loss terms, cg[j] per-lane FP64 sums of
    // the constants j < NCW used once, gcs lane j FP64 wave sums of the
    // other constant adjoints, e the error code, gxr the variable adjoints
    // of this group's points.
    auto eval_group = [&](int64_t b, uint32_t rw, uint32_t hw, int64_t base, double &lg, double (&cg)[NCW],
                          double &gcs, int32_t &e, float (&gxr)[NXR_ROWS][PPT]) {
        const uint32_t h0 = rl_u32(hw, 4);
        const int NI = (int)(h0 & 255u), nextra = (int)((h0 >> 8) & 255u), nc = (int)((h0 >> 16) & 255u);
        const int vmask = (int)(h0 >> 24);
        auto crow = [&](int j) -> int { return (int)((rl_u32(hw, j >> 2) >> (8 * (j & 3))) & 255u); };
        const int64_t yo = b * N;
        // row r of this tile group: the wave's tape, or the block's X
        auto row4 = [&](int r) -> float4 * {
            return (xblk && r >= LT) ? xb4 + (int64_t)(r - LT) * NP4 + base / 4 : tape + (size_t)r * Q * WAVE;
        };
        // operands of instruction ii (a unary instruction reads no second row)
        auto fetch = [&](uint32_t f, int ii, float (&a0)[PPT], float (&a1)[PPT]) {
            const bool bin = pf_bin(f);
            const bool c0 = pf_sc(f, 0), c1 = pf_sc(f, 1) || !bin;
            const float4 *p0 = row4(pf_row(f, 0)), *p1 = row4(pf_row(f, 1));
            const float s0 = __uint_as_float(rl_u32(rw, 32 + ii));
            const float s1 = bin ? __uint_as_float(rl_u32(rw, 48 + ii)) : 0.0f;
            float4 t0[Q], t1[Q];
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                t0[q] = c0 ? f4_zero() : p0[q * WAVE + lane];
                t1[q] = c1 ? f4_zero() : p1[q * WAVE + lane];
            }
#pragma unroll
            for (int q = 0; q < Q; ++q)
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    a0[4 * q + k] = c0 ? s0 : f4_get(t0[q], k);
                    a1[4 * q + k] = c1 ? s1 : f4_get(t1[q], k);
                }
        };
        bool act[PPT];
        int64_t n0[Q];
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            n0[q] = base + (int64_t)q * (WAVE * 4) + (int64_t)lane * 4;
#pragma unroll
            for (int k = 0; k < 4; ++k) act[4 * q + k] = n0[q] + k < N;
        }
        // upstream values (MODE 0) or targets (MODE 1): loaded now, used
        // after the forward pass
        float4 dq[Q];
#pragma unroll
        for (int q = 0; q < Q; ++q) dq[q] = need_dy ? load_quad(dy + yo, n0[q], N, vec) : f4_zero();
#pragma unroll
        for (int v = 0; v < NXR_ROWS; ++v)
#pragma unroll
            for (int k = 0; k < PPT; ++k) gxr[v][k] = 0.0f;
        if (!xblk && !one_group && !want_gx) put_var_rows(base, vmask);
        for (int r = NI; r < NI + nextra; ++r) {
            float4 *pr = row4(r);
#pragma unroll
            for (int q = 0; q < Q; ++q) pr[q * WAVE + lane] = f4_zero();
        }
        // ---- forward
        for (int ii = 0; ii < NI; ++ii) {
            const uint32_t f = rl_u32(rw, ii);
            const int op = pf_op(f);
            float a0[PPT], a1[PPT], r[PPT];
            fetch(f, ii, a0, a1);
            if (op == OP_COPY) {
#pragma unroll
                for (int k = 0; k < PPT; ++k) r[k] = a0[k];
            } else {
                fwd_ops<float, PPT>(op, a0, a1, r);
            }
            float4 *pw = row4(ii);
#pragma unroll
            for (int q = 0; q < Q; ++q)
                pw[q * WAVE + lane] = make_float4(r[4 * q], r[4 * q + 1], r[4 * q + 2], r[4 * q + 3]);
        }
        // ---- root value, loss and upstream gradient; the root's adjoint
        //      replaces its value
        {
            float4 *pr = row4(NI - 1);
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                const float4 yq = pr[q * WAVE + lane];
                if (y_out) store_quad(y_out + yo, n0[q], N, vec, yq);
                float gq[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const float yh = f4_get(yq, k), dvk = f4_get(dq[q], k);
                    float g0 = 0.0f;
                    if (a.mode == 1) {
                        const float d = yh - dvk;
                        if (act[4 * q + k]) lg += (double)d * (double)d;
                        g0 = (2.0f * d) * inv_n;
                    } else if (!a.fwd_only) {
                        g0 = dvk;
                    }
                    gq[k] = act[4 * q + k] ? g0 : 0.0f;
                }
                pr[q * WAVE + lane] = make_float4(gq[0], gq[1], gq[2], gq[3]);
            }
        }
        if (a.fwd_only) return;
        // ---- reverse sweep
        for (int ii = NI - 1; ii >= 0; --ii) {
            const uint32_t f = rl_u32(rw, ii), sw = rl_u32(rw, 16 + ii);
            const int op = pf_op(f);
            const bool binary = pf_bin(f);
            float g[PPT], a0[PPT], a1[PPT], ge0[PPT], ge1[PPT];
            {
                const float4 *pg = row4(ps_grow(sw));
#pragma unroll
                for (int q = 0; q < Q; ++q) {
                    const float4 t = pg[q * WAVE + lane];
                    g[4 * q] = t.x;
                    g[4 * q + 1] = t.y;
                    g[4 * q + 2] = t.z;
                    g[4 * q + 3] = t.w;
                }
            }
            fetch(f, ii, a0, a1);
            bwd_masked<float, PPT>(op, g, a0, a1, act, ge0, ge1, e);
            auto scatter = [&](int k, const float (&ge)[PPT]) {
                const uint32_t mode = ps_mode(sw, k);
                const int tg = ps_tgt(sw, k);
                if (mode - 1u < 3u) {  // PM_WRITE, PM_CROW, PM_ACC: a tape row
                    float4 *pt = row4(tg);
#pragma unroll
                    for (int q = 0; q < Q; ++q) {
                        const float4 o = mode == PM_ACC ? pt[q * WAVE + lane] : f4_zero();
                        pt[q * WAVE + lane] = make_float4(o.x + ge[4 * q], o.y + ge[4 * q + 1], o.z + ge[4 * q + 2],
                                                          o.w + ge[4 * q + 3]);
                    }
                } else if (mode == PM_VAR) {  // a variable: its adjoint for grad_X
                    if (want_gx) {
#pragma unroll
                        for (int v = 0; v < NXR_ROWS; ++v)
                            if (v == tg)
#pragma unroll
                                for (int k2 = 0; k2 < PPT; ++k2) gxr[v][k2] += ge[k2];
                    }
                } else if (mode == PM_CONST) {  // a constant used several times: wave sum per use
                    double gv = 0.0;
#pragma unroll
                    for (int k2 = 0; k2 < PPT; ++k2) gv += (double)ge[k2];
                    const double s = wave_sum_rows(gv, 4);
                    if (lane == tg) gcs += s;
                } else if (mode == PM_NOOP) {
#pragma unroll
                    for (int k2 = 0; k2 < PPT; ++k2)
                        if (ge[k2] != 0.0f) e = e > ERR_GRAD_NO_OP ? e : ERR_GRAD_NO_OP;
                }  // PM_NONE adds nothing
            };
            scatter(0, ge0);
            if (binary) scatter(1, ge1);
        }
        // ---- adjoints of the constants used once, from their rows, in FP64
        for (int j = 0; j < nc; ++j) {
            const int row = crow(j);
            if (row == (int)NO_CROW) continue;
            const float4 *pc = row4(row);
            double gv = 0.0;
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                const float4 t = pc[q * WAVE + lane];
                gv += (double)t.x;
                gv += (double)t.y;
                gv += (double)t.z;
                gv += (double)t.w;
            }
            if (j < NCW) {
#pragma unroll
                for (int r = 0; r < NCW; ++r)
                    if (r == j) cg[r] += gv;
            } else {
                const double s = wave_sum_rows(gv, 4);
                if (lane == j) gcs += s;
            }
        }
    };
    // ---- a program's results: loss, grad_C at its node slots, error code
    //      (gc: lane j its constant j's gradient; loss and e wave-uniform)
    auto finish = [&](int64_t b, uint32_t hw, double loss_sum, double gc, int32_t e) {
        if (a.mode == 1 && a.loss_out && lane == 0) a.loss_out[b] = loss_sum / (double)N;
        if (a.fwd_only) {
            if (lane == 0 && a.err_b) a.err_b[b] = 0;
            return;
        }
        const uint64_t cm0 = (uint64_t)rl_u32(hw, 5) | ((uint64_t)rl_u32(hw, 6) << 32);
        const uint64_t cm1 = (uint64_t)rl_u32(hw, 7) | ((uint64_t)rl_u32(hw, 8) << 32);
        // dL/dC back to node slots: constant node k takes its compact lane
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int k = lane + h * WAVE;
            const uint64_t cm = h == 0 ? cm0 : cm1;
            const uint64_t below = lane == 0 ? 0ull : (~0ull >> (WAVE - lane));
            const int o = (h == 0 ? 0 : __popcll(cm0)) + __popcll(cm & below);
            const double v0 = __shfl(gc, o & (WAVE - 1), WAVE);
            if (k < a.M) grad_C[b * a.M + k] = ((cm >> lane) & 1ull) ? (float)v0 : 0.0f;
        }
        e = wave_max_err(e);
        if (lane == 0) {
            if (e) atomicMax(&a.flags[1], e);
            if (a.err_b) a.err_b[b] = e;
        }
    };

    const int64_t count = (int64_t)*a.list_count;
    if (!want_gx) {
        // ---- one program over all its tile groups at a time, CH list
        //      positions per ticket
        constexpr int CH = fetch_chunk(LT);
        for (;;) {
            const int64_t c = wave_fetch_add_n(a.work_counter, CH);
            if (c >= count) break;
            const int64_t pe = c + CH < count ? c + CH : count;
            for (int64_t i = c; i < pe; ++i) {
                const int64_t b = (int64_t)a.list[i];
                const uint32_t *rec = a.plan + (size_t)b * PREC_WORDS;
                const uint32_t rw = rec[lane];
                const uint32_t hw = rec[64 + (lane & 15)];
                const int nc = (int)((rl_u32(hw, 4) >> 16) & 255u);
                double lacc = 0.0, gc = 0.0;
                double cacc[NCW];
#pragma unroll
                for (int q = 0; q < NCW; ++q) cacc[q] = 0.0;
                int32_t e = 0;
                float gxr[NXR_ROWS][PPT];
                for (int64_t base = 0; base < N; base += GRP) eval_group(b, rw, hw, base, lacc, cacc, gc, e, gxr);
                const double loss = a.mode == 1 ? wave_sum_rows(lacc, 4) : 0.0;
#pragma unroll
                for (int r = 0; r < NCW; ++r) {
                    if (r < nc) {
                        const double sr = wave_sum_rows(cacc[r], 4);
                        if (lane == r) gc += sr;
                    }
                }
                finish(b, hw, loss, gc, e);
            }
        }
        return;
    }
    // ---- grad_X: chunks of kc list positions, group by group
    const int64_t kc = (count + a.gx_rows - 1) / a.gx_rows;
    const int64_t ngrp = (N + GRP - 1) / GRP;
    for (;;) {
        const int64_t c = wave_fetch_add(a.work_counter);
        if (kc == 0 || c * kc >= count) break;
        const int64_t p0 = c * kc, p1 = p0 + kc < count ? p0 + kc : count;
        float *const gxrow = (float *)a.gx_partial + (a.gx_row0 + c) * gx_len;
        for (int64_t gi = 0; gi < ngrp; ++gi) {
            const int64_t base = gi * GRP;
            if (!xblk && !one_group) put_var_rows(base, (1 << n_x) - 1);
            float gxacc[NXR_ROWS][PPT];
#pragma unroll
            for (int v = 0; v < NXR_ROWS; ++v)
#pragma unroll
                for (int k = 0; k < PPT; ++k) gxacc[v][k] = 0.0f;
            for (int64_t i = p0; i < p1; ++i) {
                const int64_t b = (int64_t)a.list[i];
                const uint32_t *rec = a.plan + (size_t)b * PREC_WORDS;
                const uint32_t rw = rec[lane];
                const uint32_t hw = rec[64 + (lane & 15)];
                const int nc = (int)((rl_u32(hw, 4) >> 16) & 255u);
                double *tot = a.gx_tot + b * WIDE_TOTW;
                // this program's sums so far (loaded now, added to below):
                // lanes 0..nc-1 its constants, lane nc its loss, nc + 1 its
                // error code
                const int nt = nc + 2;
                const double prev = (gi > 0 && lane < nt) ? tot[lane] : 0.0;
                double lg = 0.0, gcs = 0.0;
                double cg[NCW];
#pragma unroll
                for (int q = 0; q < NCW; ++q) cg[q] = 0.0;
                int32_t e = 0;
                float gxr[NXR_ROWS][PPT];
                eval_group(b, rw, hw, base, lg, cg, gcs, e, gxr);
#pragma unroll
                for (int v = 0; v < NXR_ROWS; ++v)
#pragma unroll
                    for (int k = 0; k < PPT; ++k) gxacc[v][k] += gxr[v][k];
                // this group's sums, then the running ones
#pragma unroll
                for (int r = 0; r < NCW; ++r) {
                    if (r < nc) {
                        const double sr = wave_sum_rows(cg[r], 4);
                        if (lane == r) gcs += sr;
                    }
                }
                const double lsum = a.mode == 1 ? wave_sum_rows(lg, 4) : 0.0;
                const int32_t eg = wave_max_err(e);
                double run = lane == nc ? lsum : (lane == nc + 1 ? (double)eg : gcs);
                if (gi > 0) run = lane == nc + 1 ? (prev > run ? prev : run) : prev + run;
                if (gi + 1 < ngrp) {
                    if (lane < nt) tot[lane] = run;
                } else {
                    const double loss = uniform_f64(rl(run, nc));
                    const int32_t et = uniform_i32((int32_t)rl(run, nc + 1));
                    finish(b, hw, loss, run, et);
                }
            }
            // the chunk's variable adjoints of this group, written once
            // (every point of every variable: a zero where none was added)
#pragma unroll
            for (int v = 0; v < NXR_ROWS; ++v) {
                if (v >= n_x) continue;
#pragma unroll
                for (int q = 0; q < Q; ++q) {
                    const int64_t n0 = base + (int64_t)q * (WAVE * 4) + (int64_t)lane * 4;
                    store_quad(gxrow + (int64_t)v * N, n0, N, vec,
                               make_float4(gxacc[v][4 * q], gxacc[v][4 * q + 1], gxacc[v][4 * q + 2],
                                           gxacc[v][4 * q + 3]));
                }
            }
        }
    }
}

// Buckets by tape rows (count_rows_wave) and constant count nc: rows <= 8,
// <= 16, <= 32, <= 64 (all with nc <= 64), the rest.  One wave per
// expression; the run counts of the deterministic bucket lists follow in
// k_queue_count (sr_queue.h).  cost_of (nullable): the expression's share of
// its bucket's work for the static segments, rows + 2 (the per-expression
// set-up costs about two instructions).
__global__ __launch_bounds__(256) void k_classify_fused(const int64_t *__restrict__ ops, const int64_t *__restrict__ ch,
                                                        int M, int64_t B, int n_x, int8_t *bucket_of,
                                                        int32_t *cost_of) {
    const int64_t b = (int64_t)blockIdx.x * 4 + uniform_i32((int)(threadIdx.x / WAVE));
    if (b >= B) return;
    int nc = 0;
    const int nr = count_rows_wave(ops + b * M, ch + b * M * 2, M, n_x, nc);
    const int bucket = nc > 64 ? 4 : (nr <= 8 ? 0 : (nr <= 16 ? 1 : (nr <= 32 ? 2 : (nr <= 64 ? 3 : 4))));
    if (lane_id() == 0) {
        bucket_of[b] = (int8_t)bucket;
        if (cost_of) cost_of[b] = nr + 2;
    }
}

__global__ void k_fill_i32_e(int32_t *p, int32_t v, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_fill_u64_e(uint64_t *p, uint64_t v, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Sums P partial rows of length len in FP64 in two deterministic stages:
// chunk c of ROW_CHUNK rows -> part64[c][i] (row order), then the chunks in
// order -> out[i].
constexpr int ROW_CHUNK = 32;

// The partial rows of each bucket launch: [row0, row0 + cap).  A launch of
// k_fused_wide with grad_X fills only the first rows of its range (one per
// non-empty list chunk); a launch with one row per wave in the workspace
// fills the rows of the waves whose segment is not empty; a launch with one
// row per block fills all of its rows.
struct RowSel {
    int64_t row0[F_NBUCKETS], cap[F_NBUCKETS];
    int chunked[F_NBUCKETS];
    const int32_t *seg[F_NBUCKETS];  // row w holds data iff seg[w] < seg[w + 1] (nullptr: every row)
    const int32_t *counts;  // bucket list lengths (device)
};
__device__ inline int64_t rows_used(const RowSel &rs, int k) {
    if (!rs.chunked[k]) return rs.cap[k];
    const int64_t c = rs.counts[k], R = rs.cap[k];
    if (c <= 0 || R <= 0) return 0;
    const int64_t kc = (c + R - 1) / R;
    return (c + kc - 1) / kc;
}

template <typename T>
__global__ void k_sum_rows_1(const T *__restrict__ part, int64_t P, int64_t len, double *__restrict__ part64,
                             RowSel rs) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= len) return;
    const int64_t r0 = (int64_t)blockIdx.y * ROW_CHUNK;
    const int64_t r1 = r0 + ROW_CHUNK < P ? r0 + ROW_CHUNK : P;
    double s = 0.0;
    for (int k = 0; k < F_NBUCKETS; ++k) {  // the used rows in row order
        const int64_t lo = r0 > rs.row0[k] ? r0 : rs.row0[k];
        const int64_t e = rs.row0[k] + rows_used(rs, k);
        const int64_t hi = r1 < e ? r1 : e;
        if (lo >= hi) continue;
        if (rs.seg[k]) {
            // rows of empty segments hold no data: they count as zeros (the
            // FP64 sum starts at +0 and so never holds -0, and adding +0
            // changes nothing)
            const int32_t *sg = rs.seg[k] + (lo - rs.row0[k]);
            constexpr int U = 16;
            int64_t r = lo;
            for (; r < hi; r += U) {
                T v[U];
#pragma unroll
                for (int u = 0; u < U; ++u)
                    v[u] = (r + u < hi && sg[r - lo + u] < sg[r - lo + u + 1]) ? part[(r + u) * len + i] : T(0);
#pragma unroll
                for (int u = 0; u < U; ++u) s += (double)v[u];
            }
        } else {
            s = sum_rows_in_order(part + i, lo, hi, len, s);
        }
    }
    part64[(int64_t)blockIdx.y * len + i] = s;
}
// The rows are variable-major (v * N + n); grad_X is (N, n_x).
template <typename T>
__global__ void k_sum_rows_2(const double *__restrict__ part64, int64_t chunks, int64_t len, int64_t N, int n_x,
                             T *__restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= len) return;
    const double s = sum_rows_in_order(part64 + i, 0, chunks, len, 0.0);
    const int64_t v = i / N, n = i - v * N;
    out[n * n_x + v] = (T)s;
}

// Bucket kernels <LT, SL, PPT, WPB, MEM>.
template <typename T> struct FCfg;
template <> struct FCfg<float> {
    static constexpr int B0_PPT = 2;
};
template <> struct FCfg<double> {
    static constexpr int B0_PPT = 1;
};

// grad_X rows of one bucket's launch: mode 0 none, 1 wave rows in LDS (one
// partial row per block), 2 wave rows in the workspace (one per wave).
template <typename T, int WPB>
int gx_mode(bool want_gx, int64_t gx_len) {
    if (!want_gx) return 0;
    return (size_t)WPB * gx_len * sizeof(T) <= GX_LDS_ROWS_MAX ? 1 : 2;
}

template <typename T, int LT, int SL, int PPT, int WPB, bool MEM>
size_t fused_lds(int mode, int64_t gx_len) {
    return (mode == 1 ? (((size_t)WPB * gx_len * sizeof(T) + 15) / 16 * 16) : 0) +
           (MEM ? 0 : (size_t)WPB * f_tape_elems(LT, PPT) * sizeof(T));
}

// Persistent grid of one bucket: as many blocks as the CUs hold at once; with
// wave rows in the workspace, few enough that the rows stay within budget.
template <typename T, int LT, int SL, int PPT, int WPB, bool MEM>
int fused_grid(int mode, int64_t gx_len, int n_x, int64_t) {
    int per_cu = 0;
    if (MEM) {
        per_cu = 2;
    } else {
        const size_t lds = fused_lds<T, LT, SL, PPT, WPB, MEM>(mode, gx_len);
        const void *fn = n_x <= F_NX ? (const void *)k_fused_code<T, LT, SL, PPT, WPB, MEM, false>
                                     : (const void *)k_fused_code<T, LT, SL, PPT, WPB, MEM, true>;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, WPB * 64, lds) != hipSuccess || per_cu <= 0)
            per_cu = 1;
    }
    int64_t grid = (int64_t)num_cus() * per_cu;
    if (mode == 2) {
        const int64_t row_bytes = gx_len * (int64_t)sizeof(T) * WPB;
        const int64_t cap = (int64_t)GX_GLOBAL_BUDGET / (row_bytes > 0 ? row_bytes : 1);
        grid = grid < cap ? grid : (cap > 1 ? cap : 1);
    }
    return (int)grid;
}

template <typename T, int LT, int SL, int PPT, int WPB, bool MEM>
int fused_launch(const FArgs &a, int grid, int mode, hipStream_t s) {
    const size_t lds = fused_lds<T, LT, SL, PPT, WPB, MEM>(mode, a.N * a.n_x);
    if (lds > lds_limit_device()) return -3;
    if (a.n_x <= F_NX)
        hipLaunchKernelGGL((k_fused_code<T, LT, SL, PPT, WPB, MEM, false>), dim3((unsigned)grid), dim3(WPB * 64), lds,
                           s, a);
    else
        hipLaunchKernelGGL((k_fused_code<T, LT, SL, PPT, WPB, MEM, true>), dim3((unsigned)grid), dim3(WPB * 64), lds,
                           s, a);
    return (int)hipGetLastError();
}

// Launch variants of the bucket kernels (waves per block x points per lane
// per dispatch), chosen per call (sr_fused_options); variant 0 is the
// default of each bucket (resolve_variants below picks among the first two
// by N).  Only the FP32 kernels have alternatives.
struct FVar {
    int wpb, ppt;
    int (*grid)(int mode, int64_t gx_len, int n_x, int64_t N);
    int (*launch)(const FArgs &a, int grid, int mode, hipStream_t s);
    int (*gxmode)(bool want_gx, int64_t gx_len);
    bool wide;  // k_fused_wide: FP32, n_x <= NXR_ROWS, shared X, programs planned by k_plan_fused
};
template <typename T, int LT, int SL, int PPT, int WPB, bool MEM>
constexpr FVar fvar() {
    return FVar{WPB, PPT, &fused_grid<T, LT, SL, PPT, WPB, MEM>, &fused_launch<T, LT, SL, PPT, WPB, MEM>,
                &gx_mode<T, WPB>, false};
}

// The wide evaluator's LDS (the block's X copy and WPB tapes).  With grad_X
// wanted its partial rows are in the workspace (row mode 2), one per list
// chunk.
inline int64_t wide_npad(int64_t N) { return (N + WIDE_PAD - 1) / WIDE_PAD * WIDE_PAD; }
template <int LT, int PPT, int WPB>
size_t wide_lds(int n_x, int64_t N) {
    const int64_t np = wide_npad(N);
    return wide_xblk_bytes(n_x, np) + (size_t)WPB * wide_rows(LT, n_x, np) * PPT * WAVE * sizeof(float);
}
inline int wide_gx_mode(bool want_gx, int64_t) { return want_gx ? 2 : 0; }
template <int LT, int PPT, int WPB>
int wide_grid(int mode, int64_t gx_len, int n_x, int64_t N) {
    int per_cu = 0;
    const size_t lds = wide_lds<LT, PPT, WPB>(n_x, N);
    if (lds > lds_limit_device() ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)k_fused_wide<LT, PPT, WPB>, WPB * 64,
                                                      lds) != hipSuccess ||
        per_cu <= 0)
        per_cu = 1;
    int64_t grid = (int64_t)num_cus() * per_cu;
    if (mode == 2) {
        const int64_t row_bytes = gx_len * (int64_t)sizeof(float) * WPB;
        const int64_t cap = (int64_t)GX_GLOBAL_BUDGET / (row_bytes > 0 ? row_bytes : 1);
        grid = grid < cap ? grid : (cap > 1 ? cap : 1);
    }
    return (int)grid;
}
template <int LT, int PPT, int WPB>
int wide_launch(const FArgs &a, int grid, int mode, hipStream_t s) {
    const size_t lds = wide_lds<LT, PPT, WPB>(a.n_x, a.N);
    if (lds > lds_limit_device()) return -3;
    hipLaunchKernelGGL((k_fused_wide<LT, PPT, WPB>), dim3((unsigned)grid), dim3(WPB * 64), lds, s, a);
    return (int)hipGetLastError();
}
template <int LT, int PPT, int WPB>
constexpr FVar fvar_wide() {
    return FVar{WPB, PPT, &wide_grid<LT, PPT, WPB>, &wide_launch<LT, PPT, WPB>, &wide_gx_mode, true};
}

constexpr int F_MAXVAR = 6;
template <typename T> struct FTab;
template <> struct FTab<float> {
    static constexpr int n[F_NBUCKETS] = {6, 5, 3, 2, 1};
    static constexpr FVar v[F_NBUCKETS][F_MAXVAR] = {
        {fvar_wide<8, 4, 4>(), fvar_wide<8, 8, 2>(), fvar<float, 8, 1, 2, 4, false>(),
         fvar<float, 8, 1, 1, 4, false>(), fvar<float, 8, 1, 4, 4, false>(), fvar<float, 8, 1, 2, 2, false>()},
        {fvar_wide<16, 4, 2>(), fvar_wide<16, 4, 4>(), fvar<float, 16, 1, 2, 4, false>(),
         fvar<float, 16, 1, 1, 4, false>(), fvar<float, 16, 1, 2, 2, false>(), FVar{}},
        {fvar<float, 32, 1, 1, 2, false>(), fvar<float, 32, 1, 2, 2, false>(), fvar<float, 32, 1, 1, 4, false>(),
         FVar{}, FVar{}, FVar{}},
        {fvar<float, 64, 1, 1, 2, false>(), fvar<float, 64, 1, 1, 4, false>(), FVar{}, FVar{}, FVar{}, FVar{}},
        {fvar<float, 256, 2, 1, 2, true>(), FVar{}, FVar{}, FVar{}, FVar{}, FVar{}},
    };
};
template <> struct FTab<double> {
    static constexpr int n[F_NBUCKETS] = {1, 1, 1, 1, 1};
    static constexpr FVar v[F_NBUCKETS][F_MAXVAR] = {
        {fvar<double, 8, 1, 1, 4, false>(), FVar{}, FVar{}, FVar{}, FVar{}, FVar{}},
        {fvar<double, 16, 1, 1, 4, false>(), FVar{}, FVar{}, FVar{}, FVar{}, FVar{}},
        {fvar<double, 32, 1, 1, 2, false>(), FVar{}, FVar{}, FVar{}, FVar{}, FVar{}},
        {fvar<double, 64, 1, 1, 2, false>(), FVar{}, FVar{}, FVar{}, FVar{}, FVar{}},
        {fvar<double, 256, 2, 1, 2, true>(), FVar{}, FVar{}, FVar{}, FVar{}, FVar{}},
    };
};

// Workspace: lists [5][B] int32, counters [5]+[5] | queue | segment starts
// and segment queue workspace (grad_X wanted) | grad_X partial rows | FP64
// chunk sums of the rows | global tape (bucket 4) | plan records and
// transposed X (k_fused_wide) | per-program running sums (k_fused_wide with
// grad_X).
struct FWs {
    size_t q_off, seg_off, sq_off, gx_off, g64_off, gt_off, plan_off, xt_off, tot_off, total;
    int64_t nrows[F_NBUCKETS];  // partial rows per bucket
    int chunked[F_NBUCKETS];    // k_fused_wide with grad_X: rows per list chunk
    bool wide;              // some bucket runs k_fused_wide: plans and xt are laid out
    int grid[F_NBUCKETS];   // blocks per bucket launch (0: not launched)
    int mode[F_NBUCKETS];   // grad_X row mode per bucket
    int64_t row0[F_NBUCKETS];  // first partial row per bucket
    int64_t rows;           // partial rows in all
    int64_t W[F_NBUCKETS];  // waves per bucket launch (segments)
};
// var: each bucket's variant; nullptr = the largest workspace over every
// variant of every bucket (what sr_fused_workspace_bytes reports, so one
// workspace serves any choice).
template <typename T>
FWs fused_layout(int64_t N, int64_t n_x, int64_t B, int64_t M, bool want_gx, const int *var = nullptr) {
    FWs w;
    const int64_t gx_len = N * n_x;
    const int ni = (int)n_x;
    // a program of M nodes needs at most 2 M - 1 tape rows
    const bool used[F_NBUCKETS] = {true, M > 4, M > 8, M > 16, M > 32};
    int64_t rows_k[F_NBUCKETS];
    for (int k = 0; k < F_NBUCKETS; ++k) {
        w.grid[k] = 0;
        w.mode[k] = 0;
        w.W[k] = 0;
        w.chunked[k] = 0;
        rows_k[k] = 0;
        if (!used[k]) continue;
        const int v0 = var ? var[k] : 0, v1 = var ? var[k] + 1 : FTab<T>::n[k];
        for (int v = v0; v < v1; ++v) {
            const FVar &f = FTab<T>::v[k][v];
            const int mode = f.gxmode(want_gx, gx_len);
            const int grid = f.grid(mode, gx_len, ni, N);
            const int64_t W = (int64_t)grid * f.wpb;
            // k_fused_wide: WIDE_CHUNKS_PER_WAVE list chunks per wave, each
            // with its own row
            int64_t rows = mode == 1 ? grid : (mode == 2 ? W : 0);
            if (f.wide && mode == 2) {
                const int64_t cap = (int64_t)GX_GLOBAL_BUDGET / (gx_len * (int64_t)sizeof(T) > 0 ? gx_len * (int64_t)sizeof(T) : 1);
                const int64_t want = W * WIDE_CHUNKS_PER_WAVE;
                rows = want < cap ? want : (cap > W ? cap : W);
            }
            if (var) {
                w.grid[k] = grid;
                w.mode[k] = mode;
                w.chunked[k] = f.wide && mode == 2;
            } else if (grid > w.grid[k]) {
                w.grid[k] = grid;
            }
            w.W[k] = W > w.W[k] ? W : w.W[k];
            rows_k[k] = rows > rows_k[k] ? rows : rows_k[k];
        }
    }
    w.rows = 0;
    for (int k = 0; k < F_NBUCKETS; ++k) {
        w.row0[k] = w.rows;
        w.nrows[k] = rows_k[k];
        w.rows += rows_k[k];
    }
    size_t off = (((size_t)F_NBUCKETS * B * 4 + 64) + 255) / 256 * 256;
    w.q_off = off;
    off += (queue_ws_bytes(B, F_NBUCKETS) + 255) / 256 * 256;
    w.seg_off = off;
    if (want_gx)
        for (int k = 0; k < F_NBUCKETS; ++k) off += ((size_t)(w.W[k] + 1) * 4 + 255) / 256 * 256;
    w.sq_off = off;
    if (want_gx) off += (queue_seg_ws_bytes(B, F_NBUCKETS) + 255) / 256 * 256;
    w.gx_off = off;
    if (want_gx) off += (size_t)w.rows * gx_len * sizeof(T);
    off = (off + 255) / 256 * 256;
    w.g64_off = off;
    if (want_gx) off += (size_t)((w.rows + ROW_CHUNK - 1) / ROW_CHUNK) * gx_len * sizeof(double);
    off = (off + 255) / 256 * 256;
    w.gt_off = off;
    if (w.grid[4]) off += (size_t)w.grid[4] * FTab<T>::v[4][0].wpb * f_tape_elems(256, 1) * sizeof(T);
    off = (off + 255) / 256 * 256;
    // plan records and the transposed X of the wide evaluator
    w.wide = false;
    for (int k = 0; k < 2; ++k)
        for (int v = var ? var[k] : 0; v < (var ? var[k] + 1 : FTab<T>::n[k]); ++v)
            w.wide = w.wide || FTab<T>::v[k][v].wide;
    w.plan_off = off;
    w.xt_off = off;
    if (w.wide) {
        off += (size_t)B * PREC_WORDS * sizeof(uint32_t);
        off = (off + 255) / 256 * 256;
        w.xt_off = off;
        off += (size_t)NXR_ROWS * wide_npad(N) * sizeof(float);
        off = (off + 255) / 256 * 256;
    }
    w.tot_off = off;
    if (w.wide && want_gx) off += (size_t)B * WIDE_TOTW * sizeof(double);
    w.total = off + 256;
    return w;
}

// The variant of each bucket: the caller's choice, or (no options, or -1)
// the automatic one: variant 0, except that a bucket whose variant 0 takes
// two points per lane and variant 1 one switches to variant 1 when N padded
// to 64-point tiles is at most 80% of N padded to 128-point ones (at N = 64
// half of every 128-point tile is padding; two points per lane gain about
// 15% at equal work).  False for a choice out of range.
// wide_ok: the call may run k_fused_wide (FP32, at most NXR_ROWS variables,
// shared X); otherwise a wide choice counts as automatic, and the automatic
// choice skips the wide variants.
template <typename T>
bool resolve_variants(const int *variant, int64_t N, bool wide_ok, int var[F_NBUCKETS]) {
    for (int k = 0; k < F_NBUCKETS; ++k) {
        int v = variant ? variant[k] : -1;
        if (v >= FTab<T>::n[k]) return false;
        if (v >= 0 && FTab<T>::v[k][v].wide && !wide_ok) v = -1;
        if (v < 0) {
            int v0 = 0;
            while (!wide_ok && v0 + 1 < FTab<T>::n[k] && FTab<T>::v[k][v0].wide) ++v0;
            v = v0;
            const FVar &f0 = FTab<T>::v[k][v0];
            if (!f0.wide && v0 + 1 < FTab<T>::n[k] && f0.ppt == 2 && FTab<T>::v[k][v0 + 1].ppt == 1 &&
                FTab<T>::v[k][v0 + 1].wpb == f0.wpb) {
                const int64_t pad1 = (N + 63) / 64 * 64, pad2 = (N + 127) / 128 * 128;
                if (5 * pad1 <= 4 * pad2) v = v0 + 1;
            }
        }
        var[k] = v;
    }
    return true;
}

template <typename T>
int fused_root_code(const T *X, int64_t N, int64_t n_x, const int64_t *ops, const int64_t *ch, const T *C, int64_t M,
                    int64_t B, const T *dy, int mode, int fwd_only, T *y_out, double *loss_out, T *grad_C,
                    T *grad_X, int32_t *err_b, int32_t *flags, void *workspace, int64_t workspace_bytes,
                    hipStream_t s, const int64_t *x_off = nullptr, const int64_t *y_off = nullptr,
                    const int32_t *n_pts = nullptr, bool one_stream = false, const int *variant = nullptr) {
    if (M <= 0 || M > 128 || B < 0 || N <= 0 || n_x < 0 || n_x > MAX_NX_CODE + 1 || (mode != 0 && mode != 1))
        return -1;
    const bool ragged = x_off != nullptr || y_off != nullptr || n_pts != nullptr;
    if (ragged && (x_off == nullptr || y_off == nullptr || n_pts == nullptr || grad_X != nullptr)) return -1;
    // (an empty batch may come with null arrays)
    if (B > 0 && dy == nullptr && (mode == 1 || !fwd_only)) return -1;  // the kernel reads dy in these cases
    if (B > 0 && grad_C == nullptr && !fwd_only) return -1;
    const bool want_gx = grad_X != nullptr && !fwd_only;
    const bool wide_ok = sizeof(T) == 4 && n_x <= NXR_ROWS && !ragged;
    int var[F_NBUCKETS];
    if (!resolve_variants<T>(variant, N, wide_ok, var)) return -1;
    const FWs L = fused_layout<T>(N, n_x, B, M, want_gx, var);
    if (workspace_bytes < (int64_t)L.total) return -2;
    hipLaunchKernelGGL(k_fill_i32_e, dim3(1), dim3(64), 0, s, flags, 0, (int64_t)2);
    const int64_t gx_len = N * n_x;
    if (B == 0) {
        if (want_gx) {
            const int64_t nb = (gx_len + 255) / 256;
            if (sizeof(T) == 4)
                hipLaunchKernelGGL(k_fill_i32_e, dim3((unsigned)nb), dim3(256), 0, s, (int32_t *)grad_X, 0, gx_len);
            else
                hipLaunchKernelGGL(k_fill_u64_e, dim3((unsigned)nb), dim3(256), 0, s, (uint64_t *)grad_X, 0ull,
                                   gx_len);
        }
        return (int)hipGetLastError();
    }
    char *ws = (char *)workspace;
    int32_t *lists = (int32_t *)ws;
    int32_t *counters = (int32_t *)(ws + (size_t)F_NBUCKETS * B * 4);  // [0..4] counts, [5..9] work
    hipLaunchKernelGGL(k_fill_i32_e, dim3(1), dim3(64), 0, s, counters, 0, (int64_t)(2 * F_NBUCKETS));
    void *qws = ws + L.q_off;
    QueueSegs segs{};
    if (want_gx) {
        size_t o = L.seg_off;
        for (int k = 0; k < F_NBUCKETS; ++k) {
            segs.seg[k] = (int32_t *)(ws + o);
            segs.W[k] = L.W[k];
            o += ((size_t)(L.W[k] + 1) * 4 + 255) / 256 * 256;
        }
    }
    void *sqws = want_gx ? (void *)(ws + L.sq_off) : nullptr;
    const int64_t n_pad = wide_npad(N);
    if (L.wide) {
        // programs compiled once into plan records (validation and
        // classification included); X transposed for the variable rows
        const int64_t nxt = (int64_t)NXR_ROWS * n_pad;
        hipLaunchKernelGGL(k_xt_f32, dim3((unsigned)((nxt + 255) / 256)), dim3(256), 0, s, (const float *)X, N,
                           (int)n_x, n_pad, (float *)(ws + L.xt_off));
        hipLaunchKernelGGL(k_plan_fused, dim3((unsigned)((B + 3) / 4)), dim3(256), 0, s, ops, ch, (const float *)C,
                           (int)M, B, (int)n_x, queue_bucket_of(qws),
                           want_gx ? queue_cost_of(sqws) : (int32_t *)nullptr, (uint32_t *)(ws + L.plan_off), err_b,
                           flags);
    } else {
        hipLaunchKernelGGL(k_classify_fused, dim3((unsigned)((B + 3) / 4)), dim3(256), 0, s, ops, ch, (int)M, B,
                           (int)n_x, queue_bucket_of(qws), want_gx ? queue_cost_of(sqws) : (int32_t *)nullptr);
    }
    queue_finish(qws, B, F_NBUCKETS, lists, counters, s, want_gx ? &segs : nullptr, sqws);
    FArgs a;
    memset(&a, 0, sizeof(a));
    a.X = X; a.N = N; a.n_x = (int)n_x; a.ops = ops; a.ch = ch; a.C = C; a.M = (int)M; a.B = B; a.dy = dy;
    a.y_out = y_out; a.loss_out = loss_out; a.grad_C = grad_C;
    a.gx_partial = want_gx ? (void *)(ws + L.gx_off) : nullptr;
    a.err_b = err_b; a.flags = flags;
    a.gtape = L.grid[4] ? (void *)(ws + L.gt_off) : nullptr;
    a.mode = mode; a.fwd_only = fwd_only; a.want_gx = want_gx ? 1 : 0;
    a.x_off = x_off; a.y_off = y_off; a.n_pts = n_pts;
    if (L.wide) {
        a.plan = (const uint32_t *)(ws + L.plan_off);
        a.xt = (const float *)(ws + L.xt_off);
        a.n_pad = n_pad;
        a.gx_tot = want_gx ? (double *)(ws + L.tot_off) : nullptr;
        a.vec = (N % 4 == 0 && ((uintptr_t)dy & 15) == 0 && ((uintptr_t)y_out & 15) == 0) ? 1 : 0;
    }
    auto bucket_args = [&](int k) {
        FArgs r = a;
        r.list = lists + (size_t)k * B;
        r.list_count = counters + k;
        r.work_counter = counters + F_NBUCKETS + k;
        r.gx_row0 = L.row0[k];
        r.gx_rows = L.nrows[k];
        r.gx_lds = L.mode[k] == 1 ? 1 : 0;
        r.seg = want_gx ? segs.seg[k] : nullptr;
        return r;
    };
    int r = (int)hipGetLastError();
    if (r) return r;
    auto launch = [&](int k, hipStream_t st) {
        return FTab<T>::v[k][var[k]].launch(bucket_args(k), L.grid[k], L.mode[k], st);
    };
    if (M <= 4) {
        if ((r = launch(0, s))) return r;
    } else if (one_stream) {
        // every bucket on the caller's stream, one after the other (no side
        // streams, so nothing is created on first use)
        if (L.grid[4] && (r = launch(4, s))) return r;
        if (L.grid[3] && (r = launch(3, s))) return r;
        if (L.grid[2] && (r = launch(2, s))) return r;
        if (L.grid[1] && (r = launch(1, s))) return r;
        if ((r = launch(0, s))) return r;
    } else {
        // buckets 4+3, 2 and 1 on side streams, bucket 0 on the caller's
        // stream, joined back (also after a failed launch)
        ForkScope fk(s);
        if ((r = fk.fork())) return fk.finish(r);
        // buckets 4 and 3 share a stream; with M <= 64 bucket 4 only takes
        // programs with extra adjoint rows (rare) and goes second
        if (M > 64 && (r = launch(4, fk.side(0)))) return fk.finish(r);
        if (L.grid[3] && (r = launch(3, fk.side(0)))) return fk.finish(r);
        if (L.grid[4] && M <= 64 && (r = launch(4, fk.side(0)))) return fk.finish(r);
        if (L.grid[2] && (r = launch(2, fk.side(1)))) return fk.finish(r);
        if (L.grid[1] && (r = launch(1, fk.side(2)))) return fk.finish(r);
        if ((r = launch(0, s))) return fk.finish(r);
        if ((r = fk.finish(0))) return r;
    }
    if (want_gx) {
        const int64_t rows = L.rows;
        const int64_t chunks = (rows + ROW_CHUNK - 1) / ROW_CHUNK;
        const unsigned nb = (unsigned)((gx_len + 255) / 256);
        if (chunks > 65535) return -1;
        RowSel rs;
        for (int k = 0; k < F_NBUCKETS; ++k) {
            rs.row0[k] = L.row0[k];
            rs.cap[k] = L.nrows[k];
            rs.chunked[k] = L.chunked[k];
            rs.seg[k] = (L.mode[k] == 2 && !L.chunked[k]) ? segs.seg[k] : nullptr;
        }
        rs.counts = counters;
        hipLaunchKernelGGL(k_sum_rows_1<T>, dim3(nb, (unsigned)chunks), dim3(256), 0, s, (const T *)(ws + L.gx_off),
                           rows, gx_len, (double *)(ws + L.g64_off), rs);
        hipLaunchKernelGGL(k_sum_rows_2<T>, dim3(nb), dim3(256), 0, s, (const double *)(ws + L.g64_off), chunks,
                           gx_len, N, (int)n_x, grad_X);
    }
    return (int)hipGetLastError();
}

// ---------------------------------------------------------------------------
// Allocation-free form: the workspace is a file-scope device array of fixed
// size instead of caller memory.  Each GPU holds its own copy, whose device
// address is looked up once per device (hipGetSymbolAddress) and kept; the
// array's name itself is only a host-side handle and is never passed to a
// kernel.  Every call on a GPU shares that GPU's copy.
// ---------------------------------------------------------------------------
constexpr size_t FUSED_SCRATCH_BYTES = 128u << 20;
__device__ __align__(256) unsigned char scratch[FUSED_SCRATCH_BYTES];

constexpr int SCRATCH_MAX_DEVICES = 64;

// Device address of this GPU's scratch array, or nullptr when the lookup
// fails (*err receives its code).  Looked up on the first call per device.
inline void *scratch_ptr(int *err) {
    static std::mutex mu;
    static void *ptrs[SCRATCH_MAX_DEVICES] = {nullptr};
    int dev = 0;
    *err = (int)hipGetDevice(&dev);
    if (*err) return nullptr;
    if (dev < 0 || dev >= SCRATCH_MAX_DEVICES) {
        *err = -1;
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(mu);
    if (ptrs[dev] == nullptr) {
        void *p = nullptr;
        *err = (int)hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (*err) return nullptr;
        ptrs[dev] = p;
    }
    return ptrs[dev];
}

}  // namespace

extern "C" {

int64_t sr_fused_noalloc_max_bytes(void) { return (int64_t)FUSED_SCRATCH_BYTES; }

// sr_fused_root_f32 without a workspace argument (see include/sr_eval.h).
int sr_fused_root_noalloc_f32(const float *X, int64_t N, int64_t n_x, const int64_t *ops, const int64_t *ch,
                              const float *C, int64_t M, int64_t B, const float *dy, int mode, int fwd_only,
                              float *y_out, double *loss_out, float *grad_C, float *grad_X, int32_t *err_b,
                              int32_t *flags, void *stream) {
    if (M <= 0 || M > 128 || B < 0 || N <= 0 || n_x < 0) return -1;
    const bool want_gx = grad_X != nullptr && !fwd_only;
    int var[F_NBUCKETS];
    if (!resolve_variants<float>(nullptr, N, n_x <= NXR_ROWS, var)) return -1;
    if (fused_layout<float>(N, n_x, B, M, want_gx, var).total > FUSED_SCRATCH_BYTES) return -2;
    int e = 0;
    void *ws = scratch_ptr(&e);
    if (ws == nullptr) return e ? e : -1;
    return fused_root_code<float>(X, N, n_x, ops, ch, C, M, B, dy, mode, fwd_only, y_out, loss_out, grad_C, grad_X,
                                  err_b, flags, ws, (int64_t)FUSED_SCRATCH_BYTES, (hipStream_t)stream, nullptr,
                                  nullptr, nullptr, true);
}


int64_t sr_fused_workspace_bytes(int64_t N, int64_t n_x, int64_t B, int64_t M, int want_gx, int is_f64) {
    return is_f64 ? (int64_t)fused_layout<double>(N, n_x, B, M, want_gx != 0).total
                  : (int64_t)fused_layout<float>(N, n_x, B, M, want_gx != 0).total;
}

int sr_fused_root_f32(const float *X, int64_t N, int64_t n_x, const int64_t *ops, const int64_t *ch, const float *C,
                      int64_t M, int64_t B, const float *dy, int mode, int fwd_only, float *y_out, double *loss_out,
                      float *grad_C, float *grad_X, int32_t *err_b, int32_t *flags, void *workspace,
                      int64_t workspace_bytes, void *stream) {
    return fused_root_code<float>(X, N, n_x, ops, ch, C, M, B, dy, mode, fwd_only, y_out, loss_out, grad_C, grad_X,
                                  err_b, flags, workspace, workspace_bytes, (hipStream_t)stream);
}
int sr_fused_root_f64(const double *X, int64_t N, int64_t n_x, const int64_t *ops, const int64_t *ch,
                      const double *C, int64_t M, int64_t B, const double *dy, int mode, int fwd_only,
                      double *y_out, double *loss_out, double *grad_C, double *grad_X, int32_t *err_b,
                      int32_t *flags, void *workspace, int64_t workspace_bytes, void *stream) {
    return fused_root_code<double>(X, N, n_x, ops, ch, C, M, B, dy, mode, fwd_only, y_out, loss_out, grad_C, grad_X,
                                   err_b, flags, workspace, workspace_bytes, (hipStream_t)stream);
}

// Launch-variant table of the fused evaluator and its per-call form.
int sr_fused_num_variants(int bucket, int is_f64) {
    if (bucket < 0 || bucket >= F_NBUCKETS) return -1;
    return is_f64 ? FTab<double>::n[bucket] : FTab<float>::n[bucket];
}
int sr_fused_variant_info(int bucket, int variant, int is_f64, int *waves_per_block, int *ppt) {
    if (bucket < 0 || bucket >= F_NBUCKETS || variant < 0 || variant >= sr_fused_num_variants(bucket, is_f64))
        return -1;
    const FVar &f = is_f64 ? FTab<double>::v[bucket][variant] : FTab<float>::v[bucket][variant];
    if (waves_per_block) *waves_per_block = f.wpb;
    if (ppt) *ppt = f.ppt;
    return 0;
}
int sr_fused_root_opts_f32(const float *X, int64_t N, int64_t n_x, const int64_t *ops, const int64_t *ch,
                           const float *C, int64_t M, int64_t B, const float *dy, int mode, int fwd_only,
                           float *y_out, double *loss_out, float *grad_C, float *grad_X, int32_t *err_b,
                           int32_t *flags, const sr_fused_options *opts, void *workspace, int64_t workspace_bytes,
                           void *stream) {
    return fused_root_code<float>(X, N, n_x, ops, ch, C, M, B, dy, mode, fwd_only, y_out, loss_out, grad_C, grad_X,
                                  err_b, flags, workspace, workspace_bytes, (hipStream_t)stream, nullptr, nullptr,
                                  nullptr, false, opts ? opts->variant : nullptr);
}
int sr_fused_root_opts_f64(const double *X, int64_t N, int64_t n_x, const int64_t *ops, const int64_t *ch,
                           const double *C, int64_t M, int64_t B, const double *dy, int mode, int fwd_only,
                           double *y_out, double *loss_out, double *grad_C, double *grad_X, int32_t *err_b,
                           int32_t *flags, const sr_fused_options *opts, void *workspace, int64_t workspace_bytes,
                           void *stream) {
    return fused_root_code<double>(X, N, n_x, ops, ch, C, M, B, dy, mode, fwd_only, y_out, loss_out, grad_C, grad_X,
                                   err_b, flags, workspace, workspace_bytes, (hipStream_t)stream, nullptr, nullptr,
                                   nullptr, false, opts ? opts->variant : nullptr);
}

// Ragged: expression b has n_pts[b] points at X + x_off[b] (row-major,
// n_x per row) and its dy / y values at offset y_off[b]; N is the largest
// n_pts (only its bound matters).  No grad_X.
int sr_fused_root_ragged_f32(const float *X, int64_t N, int64_t n_x, const int64_t *x_off, const int64_t *y_off,
                             const int32_t *n_pts, const int64_t *ops, const int64_t *ch, const float *C, int64_t M,
                             int64_t B, const float *dy, int mode, int fwd_only, float *y_out, double *loss_out,
                             float *grad_C, int32_t *err_b, int32_t *flags, void *workspace, int64_t workspace_bytes,
                             void *stream) {
    if (B > 0 && (x_off == nullptr || y_off == nullptr || n_pts == nullptr)) return -1;
    return fused_root_code<float>(X, N, n_x, ops, ch, C, M, B, dy, mode, fwd_only, y_out, loss_out, grad_C, nullptr,
                                  err_b, flags, workspace, workspace_bytes, (hipStream_t)stream, x_off, y_off, n_pts);
}
int sr_fused_root_ragged_f64(const double *X, int64_t N, int64_t n_x, const int64_t *x_off, const int64_t *y_off,
                             const int32_t *n_pts, const int64_t *ops, const int64_t *ch, const double *C, int64_t M,
                             int64_t B, const double *dy, int mode, int fwd_only, double *y_out, double *loss_out,
                             double *grad_C, int32_t *err_b, int32_t *flags, void *workspace, int64_t workspace_bytes,
                             void *stream) {
    if (B > 0 && (x_off == nullptr || y_off == nullptr || n_pts == nullptr)) return -1;
    return fused_root_code<double>(X, N, n_x, ops, ch, C, M, B, dy, mode, fwd_only, y_out, loss_out, grad_C, nullptr,
                                   err_b, flags, workspace, workspace_bytes, (hipStream_t)stream, x_off, y_off,
                                   n_pts);
}

}  // extern "C"
