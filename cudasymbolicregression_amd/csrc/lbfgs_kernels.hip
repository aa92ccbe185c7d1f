// Device-resident batched L-BFGS for expression constants.
//
// One wavefront owns one fitting problem for its whole optimisation: the
// expression program, the L-BFGS state (x, g, the m-pair history, the line
// search state) and the convergence test never leave the CU.  Lane d holds
// constant slot d (and d + 64 when M > 64), so every L-BFGS vector operation
// is a per-lane VALU op and every inner product is one FP64 wave reduction.
// Each objective evaluation ("closure") is the fused root-only forward +
// backward of sr_wave.h over the problem's points: loss and dL/dC are reduced
// in FP64 inside the wave.
//
// Algorithm and options follow the reference's LBFGS_batch.step with
// optimize's closure (symbolic_torch/lbfgs_batch.py:65-326,
// symbolic_torch/optimization.py:63-101), per problem:
//   * p = -H g by the two-loop recursion (rho = 1/(y.s + eps),
//     gamma = s.y/(y.y + eps) from the newest pair, no scaling before the
//     first pair);
//   * strong Wolfe bracketing (c1 = 1e-4, c2 = 0.9, alpha0 = 1,
//     alpha <- alpha + 0.8*(20 - alpha), acceptance on the curvature test)
//     with the bisection zoom, or Armijo backtracking (halving);
//   * x <- x + alpha p, history push (drop oldest beyond m);
//   * stop when ||g||_2 < tol_grad (NaN never stops).
// Closures the reference recomputes at points it already evaluated (phi_0 at
// x, phi_lo, and f at the accepted step) are reused: evaluation is
// deterministic, so the values are identical.  The inner products, the norm,
// the losses and the Wolfe tests run in FP64; vectors and step lengths stay
// FP32 like the reference's tensors.  A problem stops early when it has
// converged, when its loss reaches the optional per-problem stop_loss, or
// when every slot of x is NaN with a non-empty history (the reference's
// iterates stay NaN from there on).
#include "sr_wave.h"
#include "sr_device_info.h"
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

using namespace sr;

namespace {

constexpr int MHMAX = 16;  // largest supported history_size

struct LbfgsArgs {
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
    int32_t *queue;
    int32_t *queue_count;
    int32_t *work_counter;
    const int32_t *list;
    const int32_t *list_count;
    float *gtape;            // global tape/adjoint slots for programs too long for LDS
};

template <int SL>
struct Vec {
    float v[SL];
};

// Inner product over the M slots (lanes past M never contribute).
template <int SL>
__device__ inline double dot(const Vec<SL> &a, const Vec<SL> &b, int M) {
    const int lane = lane_id();
    double s = 0.0;
#pragma unroll
    for (int h = 0; h < SL; ++h)
        if (lane + h * WAVE < M) s += (double)a.v[h] * (double)b.v[h];
    return wave_sum(s);
}

template <int SL>
__device__ inline bool all_finite(const Vec<SL> &a, int M) {
    const int lane = lane_id();
    bool bad = false;
#pragma unroll
    for (int h = 0; h < SL; ++h)
        if (lane + h * WAVE < M && !isfinite(a.v[h])) bad = true;
    return __ballot(bad) == 0ull;
}

template <int SL>
__device__ inline bool all_nan(const Vec<SL> &a, int M) {
    const int lane = lane_id();
    bool ok = true;
#pragma unroll
    for (int h = 0; h < SL; ++h) {
        int k = lane + h * WAVE;
        if (k < M && !(a.v[h] != a.v[h])) ok = false;
    }
    return __ballot(!ok) == 0ull;
}

// One objective evaluation at the constants held in p.c0/p.c1.
template <int LMAX, int SL>
__device__ inline double closure(Program<float> &p, const float *X, const float *Yb, int64_t N, int n_x,
                                 float *tape, float *adj, double *cextra, Vec<SL> &g, int32_t &err) {
    const int lane = lane_id();
    const int L = p.L;
    for (int k = lane; k < LMAX; k += WAVE) cextra[k] = 0.0;
    GradAcc<float> acc;
#pragma unroll
    for (int q = 0; q < NCREG; ++q) acc.cacc[q] = 0.0;
    double loss = 0.0;
    int32_t e = 0;
    const float inv_n = 1.0f / (float)N;
    const int64_t n_tiles = (N + WAVE - 1) / WAVE;
    for (int64_t t = 0; t < n_tiles; ++t) {
        const int64_t n = t * WAVE + lane;
        const bool active = n < N;
        PointCtx<float> pc = load_point<float>(X, n, active, n_x);
        forward_tile<float>(p, pc, tape, L);
        float yh = L > 0 ? tape[p.root * WAVE + lane] : __builtin_nanf("");
        float target = active ? Yb[n] : 0.0f;
        float d = yh - target;
        if (active) loss += (double)d * (double)d;
        float g0 = (2.0f * d) * inv_n;
        if (L == 0) {
            if (active && g0 != 0.0f) e = ERR_GRAD_NO_OP;
            continue;
        }
        for (int k = 0; k < L; ++k) adj[k * WAVE + lane] = 0.0f;
        adj[p.root * WAVE + lane] = active ? g0 : 0.0f;
        auto val = [&](int k) -> float { return tape[k * WAVE + lane]; };
        auto adjr = [&](int k) -> float & { return adj[k * WAVE + lane]; };
        int32_t ee = backward_tile<float, false>(p, val, adjr, L, active, acc, cextra, n_x, (float *)nullptr);
        e = ee > e ? ee : e;
    }
    double lo, hi;
    finish_constants<float>(p, acc, cextra, lo, hi);
    g.v[0] = (float)lo;
    if (SL > 1) g.v[SL - 1] = (float)hi;
    err = wave_max_i32(e);
    return wave_sum(loss) / (double)N;
}

template <int SL>
__device__ inline void set_point(Program<float> &p, const Vec<SL> &x, const Vec<SL> &d, float a) {
    // trial constants x + a*d, rounded like the reference's x + alpha*p
    p.c0 = x.v[0] + a * d.v[0];
    if (SL > 1) p.c1 = x.v[SL - 1] + a * d.v[SL - 1];
}

template <int LMAX, int SL, int W, bool GT>
__global__ __launch_bounds__(W * 64) void k_lbfgs(LbfgsArgs a) {
    // GT: tape and adjoints in a global workspace slot instead of LDS.
    extern __shared__ __align__(16) unsigned char smem[];
    const int wave = uniform_i32((int)(threadIdx.x / WAVE));
    const int lane = lane_id();
    constexpr size_t TAPE_LDS = GT ? 0 : (size_t)W * 2 * LMAX * WAVE;
    float *tape = GT ? a.gtape + ((size_t)blockIdx.x * W + wave) * 2 * LMAX * WAVE
                     : reinterpret_cast<float *>(smem) + (size_t)wave * 2 * LMAX * WAVE;
    float *adj = tape + LMAX * WAVE;
    double *cextra = reinterpret_cast<double *>(reinterpret_cast<float *>(smem) + TAPE_LDS) + (size_t)wave * LMAX;
    const int64_t count = a.list ? (int64_t)*a.list_count : a.B;
    const double c1 = 1e-4, c2 = 0.9;
    for (;;) {
        int got = 0;
        if (lane == 0) got = atomicAdd(a.work_counter, 1);
        const int i = uniform_i32(got);  // wave-uniform slot index (SGPR)
        if (i >= count) break;
        const int64_t b = a.list ? (int64_t)a.list[i] : (int64_t)i;
        Program<float> p = load_program<float>(a.ops, a.ch, a.C_init, b, a.M, 1, a.M, 1, a.M, a.n_x);
        if (p.verr) {
            if (lane == 0) {
                atomicMax(&a.flags[0], p.verr);
                a.err_out[b] = p.verr;
                a.iters_out[b] = 0;
                a.loss_out[b] = __builtin_nan("");
            }
            continue;
        }
        if (p.L > LMAX) {
            if (lane == 0) {
                int slot = atomicAdd(a.queue_count, 1);
                a.queue[slot] = (int32_t)b;
            }
            continue;
        }
        const float *X = a.X + (a.x_off ? a.x_off[b] : 0);
        const float *Yb = a.Y + (a.y_off ? a.y_off[b] : b * a.N);
        const int64_t N = a.n_pts ? (int64_t)a.n_pts[b] : a.N;
        const double stop = a.stop_loss ? a.stop_loss[b] : -1.0;

        Vec<SL> x, g, d;
        x.v[0] = p.c0;
        if (SL > 1) x.v[SL - 1] = p.c1;
        Vec<SL> S[MHMAX], Yh[MHMAX];
        double rho[MHMAX], alph[MHMAX];
        int nh = 0;
        int32_t err = 0;
        int nclos = 1;
        double f = closure<LMAX, SL>(p, X, Yb, N, a.n_x, tape, adj, cextra, g, err);
        int it = 0;
        for (; it < a.max_iter && err == 0; ++it) {
            const double gn = sqrt(dot<SL>(g, g, a.M));
            if (gn < a.tol_grad) break;
            if (f <= stop) break;
            if (nh > 0 && all_nan<SL>(x, a.M)) break;
            // ---- two-loop recursion: d = -H g
            Vec<SL> q = g;
#pragma unroll
            for (int j = MHMAX - 1; j >= 0; --j) {
                if (j < nh) {
                    alph[j] = rho[j] * dot<SL>(S[j], q, a.M);
                    const float aj = (float)alph[j];
#pragma unroll
                    for (int h = 0; h < SL; ++h) q.v[h] = q.v[h] - aj * Yh[j].v[h];
                }
            }
            Vec<SL> z = q;
            if (nh > 0) {
                Vec<SL> sl = S[0], yl = Yh[0];
#pragma unroll
                for (int j = 0; j < MHMAX; ++j)
                    if (j == nh - 1) { sl = S[j]; yl = Yh[j]; }
                const float gamma = (float)(dot<SL>(sl, yl, a.M) / (dot<SL>(yl, yl, a.M) + a.eps));
#pragma unroll
                for (int h = 0; h < SL; ++h) z.v[h] = gamma * q.v[h];
#pragma unroll
                for (int j = 0; j < MHMAX; ++j) {
                    if (j < nh) {
                        const double beta = rho[j] * dot<SL>(Yh[j], z, a.M);
                        const float coef = (float)(alph[j] - beta);
#pragma unroll
                        for (int h = 0; h < SL; ++h) z.v[h] = z.v[h] + S[j].v[h] * coef;
                    }
                }
            }
#pragma unroll
            for (int h = 0; h < SL; ++h) d.v[h] = -z.v[h];

            // ---- line search; on exit (alpha, f_new, g_new) describe x + alpha d
            const double dphi0 = dot<SL>(g, d, a.M);
            // x + 0*d equals x only for a finite direction; otherwise the
            // reference's evaluation at alpha = 0 is not the one at x.
            const bool d_finite = all_finite<SL>(d, a.M);
            float alpha = 0.0f;
            double f_new = 0.0;
            Vec<SL> g_new = g;
            bool have = false;
            int32_t e2 = 0;
            if (a.line_search == 0) {
                float a_prev = 0.0f, ai = 1.0f;
                double phi_prev = f;
                Vec<SL> g_prev = g;
                bool settled = false;
                for (int ls = 0; ls < a.ls_max_iter && !settled; ++ls) {
                    Vec<SL> ga;
                    set_point<SL>(p, x, d, ai);
                    const double phi = closure<LMAX, SL>(p, X, Yb, N, a.n_x, tape, adj, cextra, ga, e2);
                    ++nclos;
                    if (e2) break;
                    const double dphi = dot<SL>(ga, d, a.M);
                    const bool bad = (phi > f + c1 * (double)ai * dphi0) || (ls > 0 && phi >= phi_prev);
                    const bool good = fabs(dphi) <= c2 * fabs(dphi0);
                    const bool uphill = dphi >= 0.0;
                    if (good) {
                        alpha = ai; f_new = phi; g_new = ga; have = true; settled = true;
                        break;
                    }
                    if (bad || uphill) {
                        float lo = bad ? a_prev : ai, hi = bad ? ai : a_prev;
                        double phi_lo = bad ? phi_prev : phi;
                        Vec<SL> g_lo = bad ? g_prev : ga;
                        if (bad && ls == 0 && !d_finite) {
                            set_point<SL>(p, x, d, 0.0f);
                            phi_lo = closure<LMAX, SL>(p, X, Yb, N, a.n_x, tape, adj, cextra, g_lo, e2);
                            ++nclos;
                            if (e2) break;
                        }
                        bool zdone = false;
                        for (int zi = 0; zi < 10; ++zi) {
                            const float aj = (lo + hi) / 2.0f;
                            Vec<SL> gj;
                            set_point<SL>(p, x, d, aj);
                            const double phij = closure<LMAX, SL>(p, X, Yb, N, a.n_x, tape, adj, cextra, gj, e2);
                            ++nclos;
                            if (e2) break;
                            const double dphij = dot<SL>(gj, d, a.M);
                            const bool zbad = (phij > f + c1 * (double)aj * dphi0) || (phij >= phi_lo);
                            if (fabs(dphij) <= c2 * fabs(dphi0)) {
                                alpha = aj; f_new = phij; g_new = gj; have = true; zdone = true;
                                break;
                            }
                            const float new_hi = zbad ? aj : hi;
                            const bool flip = (double)dphij * (double)(new_hi - lo) >= 0.0;
                            hi = (!zbad && flip) ? lo : new_hi;
                            if (!zbad) { lo = aj; phi_lo = phij; g_lo = gj; }
                            if (fabsf(hi - lo) < 1e-9f) {
                                alpha = aj; f_new = phij; g_new = gj; have = true; zdone = true;
                                break;
                            }
                        }
                        if (e2) break;
                        if (!zdone) { alpha = lo; f_new = phi_lo; g_new = g_lo; have = true; }
                        settled = true;
                        break;
                    }
                    a_prev = ai;
                    phi_prev = phi;
                    g_prev = ga;
                    ai = ai + (20.0f - ai) * 0.8f;
                }
                if (!settled && e2 == 0) alpha = ai;
            } else {
                float ai = 1.0f;
                for (int ls = 0; ls < a.ls_max_iter; ++ls) {
                    Vec<SL> ga;
                    set_point<SL>(p, x, d, ai);
                    const double ft = closure<LMAX, SL>(p, X, Yb, N, a.n_x, tape, adj, cextra, ga, e2);
                    ++nclos;
                    if (e2) break;
                    if (ft <= f + 1e-4 * (double)ai * dphi0) {
                        alpha = ai; f_new = ft; g_new = ga; have = true;
                        break;
                    }
                    ai = ai * 0.5f;
                }
                if (!have && e2 == 0) alpha = ai;
            }
            if (e2) { err = e2; break; }
            // ---- step and history
            Vec<SL> s;
#pragma unroll
            for (int h = 0; h < SL; ++h) {
                s.v[h] = alpha * d.v[h];
                x.v[h] = x.v[h] + s.v[h];
            }
            if (!have) {
                p.c0 = x.v[0];
                if (SL > 1) p.c1 = x.v[SL - 1];
                f_new = closure<LMAX, SL>(p, X, Yb, N, a.n_x, tape, adj, cextra, g_new, e2);
                ++nclos;
                if (e2) { err = e2; break; }
            }
            Vec<SL> yv;
#pragma unroll
            for (int h = 0; h < SL; ++h) yv.v[h] = g_new.v[h] - g.v[h];
            const double rho_new = 1.0 / (dot<SL>(yv, s, a.M) + a.eps);
            if (nh == a.history) {
#pragma unroll
                for (int j = 0; j < MHMAX - 1; ++j) {
                    if (j + 1 < nh) { S[j] = S[j + 1]; Yh[j] = Yh[j + 1]; rho[j] = rho[j + 1]; }
                }
                --nh;
            }
#pragma unroll
            for (int j = 0; j < MHMAX; ++j)
                if (j == nh) { S[j] = s; Yh[j] = yv; rho[j] = rho_new; }
            ++nh;
            g = g_new;
            f = f_new;
        }
        // ---- write back (x holds all M slots, non-constant slots included)
        if (lane < a.M) a.C_out[b * a.M + lane] = x.v[0];
        if (SL > 1 && lane + WAVE < a.M) a.C_out[b * a.M + lane + WAVE] = x.v[SL - 1];
        if (lane == 0) {
            a.loss_out[b] = f;
            a.iters_out[b] = it;
            if (a.nclos_out) a.nclos_out[b] = nclos;
            a.err_out[b] = err;
            if (err) atomicMax(&a.flags[1], err);
        }
    }
}

__global__ void k_fill_i32_l(int32_t *p, int32_t v, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

inline size_t lbfgs_lds(int LMAX, int W, bool gt) {
    return (gt ? 0 : (size_t)W * 2 * LMAX * WAVE * sizeof(float)) + (size_t)W * LMAX * sizeof(double);
}

template <int LMAX, int SL, int W>
int launch_lbfgs(const LbfgsArgs &a, int64_t grid, hipStream_t s) {
    const bool gt = lbfgs_lds(LMAX, W, false) > lds_limit();
    const size_t lds = lbfgs_lds(LMAX, W, gt);
    if (lds > lds_limit()) return -3;
    if (gt) {
        if (a.gtape == nullptr) return -2;
        hipLaunchKernelGGL((k_lbfgs<LMAX, SL, W, true>), dim3((unsigned)grid), dim3(W * 64), lds, s, a);
    } else {
        hipLaunchKernelGGL((k_lbfgs<LMAX, SL, W, false>), dim3((unsigned)grid), dim3(W * 64), lds, s, a);
    }
    return (int)hipGetLastError();
}

constexpr int P2_BLOCKS_PER_CU = 8;

inline int p2_lmax(int64_t M) { return M <= 64 ? 64 : 128; }

// Workspace: [queue: B int32][4 counters] | global tape slots (phase 2, M > 64 only).
inline size_t lbfgs_gt_off(int64_t B) { return (((size_t)B * 4 + 16) + 255) / 256 * 256; }
inline bool lbfgs_needs_gt(int64_t M) { return M > 16 && lbfgs_lds(p2_lmax(M), 1, false) > lds_limit(); }
inline int64_t lbfgs_ws(int64_t B, int64_t M) {
    size_t off = lbfgs_gt_off(B);
    if (lbfgs_needs_gt(M))
        off += (size_t)num_cus() * P2_BLOCKS_PER_CU * 2 * p2_lmax(M) * WAVE * sizeof(float);
    return (int64_t)off + 256;
}

}  // namespace

extern "C" {

int64_t sr_lbfgs_workspace_bytes(int64_t B, int64_t M) { return lbfgs_ws(B, M); }

// Fits the LEARNABLE_CONSTANT slots of B batch-first expressions to their
// targets.  See include/sr_eval.h for the argument contract.
int sr_lbfgs_fit(const float *X, const float *Y, int64_t N, int64_t n_x, const int64_t *ops, const int64_t *ch,
                 const float *C_init, int64_t M, int64_t B, const int64_t *x_off, const int64_t *y_off,
                 const int32_t *n_pts, const double *stop_loss, float *C_out, double *loss_out, int32_t *iters_out,
                 int32_t *nclos_out, int32_t *err_out, int32_t *flags, int max_iter, int history, double tol_grad, int line_search,
                 int ls_max_iter, double eps, void *workspace, int64_t ws_bytes, void *stream) {
    hipStream_t s = (hipStream_t)stream;
    if (M <= 0 || M > 128 || B < 0 || n_x < 0 || history < 1 || history > MHMAX || max_iter < 0 ||
        ls_max_iter < 0 || (line_search != 0 && line_search != 1))
        return -1;
    if (n_pts == nullptr && N <= 0) return -1;
    if (ws_bytes < sr_lbfgs_workspace_bytes(B, M)) return -2;
    hipLaunchKernelGGL(k_fill_i32_l, dim3(1), dim3(64), 0, s, flags, 0, (int64_t)2);
    if (B == 0) return (int)hipGetLastError();
    char *ws = (char *)workspace;
    int32_t *queue = (int32_t *)ws;
    int32_t *counters = (int32_t *)(ws + (size_t)B * 4);  // [0] queue count, [1] work p1, [2] work p2
    hipLaunchKernelGGL(k_fill_i32_l, dim3(1), dim3(64), 0, s, counters, 0, (int64_t)4);
    LbfgsArgs a;
    memset(&a, 0, sizeof(a));
    a.X = X; a.Y = Y; a.N = N; a.n_x = (int)n_x; a.ops = ops; a.ch = ch; a.C_init = C_init; a.M = (int)M; a.B = B;
    a.x_off = x_off; a.y_off = y_off; a.n_pts = n_pts; a.stop_loss = stop_loss;
    a.C_out = C_out; a.loss_out = loss_out; a.iters_out = iters_out; a.nclos_out = nclos_out; a.err_out = err_out;
    a.flags = flags;
    a.max_iter = max_iter; a.history = history; a.tol_grad = tol_grad; a.line_search = line_search;
    a.ls_max_iter = ls_max_iter; a.eps = eps;
    a.queue = queue; a.queue_count = counters; a.work_counter = counters + 1;
    a.list = nullptr; a.list_count = nullptr;
    a.gtape = lbfgs_needs_gt(M) ? (float *)(ws + lbfgs_gt_off(B)) : nullptr;
    const int64_t g1 = (int64_t)num_cus() * 4;
    int r = (M <= 64) ? launch_lbfgs<16, 1, 4>(a, g1, s) : launch_lbfgs<16, 2, 4>(a, g1, s);
    if (r) return r;
    if (M > 16) {
        LbfgsArgs a2 = a;
        a2.list = queue;
        a2.list_count = counters;
        a2.work_counter = counters + 2;
        const int64_t g2 = (int64_t)num_cus() * P2_BLOCKS_PER_CU;
        if (M <= 64) r = launch_lbfgs<64, 1, 1>(a2, g2, s);
        else r = launch_lbfgs<128, 2, 1>(a2, g2, s);
        if (r) return r;
    }
    return (int)hipGetLastError();
}

}  // extern "C"
