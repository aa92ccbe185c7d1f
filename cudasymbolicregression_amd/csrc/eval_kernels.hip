// gfx950 evaluation kernels and their C-ABI entry points.
//
//  * validation        (reference: src/symbolic_evaluation.cpp:7-37,
//                       src/evaluation_kernels_cpu.cpp:8-52)
//  * full-cache forward (reference: SymbolicEvaluation::forward,
//                       src/symbolic_evaluation.cpp:39-92 and the per-node
//                       forward step, src/evaluation_kernels_cpu.cpp:55-219):
//                       one launch for all M nodes instead of M launches.
//  * full-cache backward (reference: SymbolicEvaluation::backward,
//                       src/symbolic_evaluation.cpp:94-168 and
//                       src/evaluation_kernels_cpu.cpp:222-662): one launch,
//                       deterministic reductions (no float atomics).
//  * fused root-only forward+backward with the MSE closure of
//                       symbolic_torch/optimization.py:72-99 folded in: the
//                       per-expression loss and dL/dC reduce in FP64 inside
//                       the kernel, the (M,N,B) cache is never materialised.
//
// Layout conventions: the backend ("mb") entry points take the reference's
// node-first tensors Ops (M,B), Ch (M,B,2), C (M,B), cache (M,N,B).  The fused
// entry points take batch-first Ops (B,M), Ch (B,M,2), C (B,M), Y/dy (B,N).
#include "sr_wave.h"
#include "sr_device_info.h"
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

using namespace sr;

namespace {

constexpr int TILE = WAVE;

__global__ void k_fill_i32(int32_t *p, int32_t v, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_fill_u64(uint64_t *p, uint64_t v, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------------
// Validation over the node-first layout.  d_first receives
// min over failing (m,b) of ((m*B + b) << 10 | code): the code of the first
// failing node in row-major (m, b) order.
// ---------------------------------------------------------------------------
__global__ void k_validate_mb(const int64_t *__restrict__ ops, const int64_t *__restrict__ ch, int64_t M,
                              int64_t B, unsigned long long *d_first) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M * B) return;
    int64_t m = i / B;
    int32_t e = validate_node(ops[i], ch[2 * i], ch[2 * i + 1], (int)(m > 0x7fffffff ? 0x7fffffff : m));
    if (e) atomicMin(d_first, ((unsigned long long)i << 10) | (unsigned long long)e);
}

// ---------------------------------------------------------------------------
// Full-cache forward: one wave per (expression b, 64-point tile).  Block ids
// are remapped so that each XCD owns a contiguous range of (tile, b) pairs:
// the waves resident on one XCD then write neighbouring b of the same cache
// lines, which merge in that XCD's L2 before going to HBM.
// ---------------------------------------------------------------------------
__device__ inline int64_t xcd_remap(int64_t bid, int64_t nblk) {
    constexpr int NXCD = 8;
    if (nblk % NXCD) return bid;
    return (bid % NXCD) * (nblk / NXCD) + bid / NXCD;
}

template <typename T>
__global__ __launch_bounds__(64) void k_full_forward(const T *__restrict__ X, int64_t N, int n_x,
                                                     const int64_t *__restrict__ ops,
                                                     const int64_t *__restrict__ ch, const T *__restrict__ C,
                                                     int M, int64_t B, int64_t K, int64_t n_tiles,
                                                     T *__restrict__ cache) {
    // One wave per (tile, expression e = b*K + j): constant set j of expression b.
    extern __shared__ __align__(16) unsigned char smem[];
    T *tape = reinterpret_cast<T *>(smem);
    const int64_t E = B * K;
    int64_t id = xcd_remap(blockIdx.x, (int64_t)gridDim.x);
    int64_t tile = id / E, e = id % E;
    if (tile >= n_tiles) return;
    const int64_t b = e / K, j = e % K;
    const int lane = lane_id();
    Program<T> p = load_program<T>(ops, ch, C + j, b, 1, B, K, B * K, M, n_x);
    int64_t n = tile * TILE + lane;
    bool active = n < N;
    PointCtx<T> pc = load_point<T>(X, n, active, n_x);
    forward_tile<T>(p, pc, tape, p.L);
    if (!active) return;
    for (int k = 0; k < p.L; ++k) cache[(((int64_t)k * N + n) * B + b) * K + j] = tape[k * WAVE + lane];
    const T nan = Mth<T>::qnan();
    for (int k = p.L; k < M; ++k) cache[(((int64_t)k * N + n) * B + b) * K + j] = nan;
}

// ---------------------------------------------------------------------------
// Full-cache backward.  Block = 4 waves = 4 consecutive expressions, looping
// over all point tiles together.  Children values come from the saved cache;
// the adjoints are accumulated in place in grad_cache, the caller's copy of
// the upstream gradient (the reference's grad_cache, symbolic_evaluation.cpp:
// 114-115): every element is owned by one lane, so no atomics are needed.
// grad_C is reduced per wave in FP64 over all points; grad_X is summed over
// the block's 4 expressions in a fixed order per tile and written as a
// per-block partial, which k_reduce_partials sums in FP64.
// ---------------------------------------------------------------------------
constexpr int BWD_WAVES = 4;

template <typename T>
__global__ __launch_bounds__(BWD_WAVES * 64) void k_full_backward(
    T *__restrict__ grad_cache, const T *__restrict__ cache, const int64_t *__restrict__ ops,
    const int64_t *__restrict__ ch, int M, int64_t B, int64_t K, int64_t N, int n_x, T *__restrict__ grad_C,
    T *__restrict__ gx_partial, T *__restrict__ grad_X_direct, int32_t *err_flag) {
    // Wave w of block i handles expression e = 4*i + w, i.e. constant set
    // j = e % K of expression b = e / K (K = 1 for the single-set layout).
    extern __shared__ __align__(16) unsigned char smem[];
    const int wave = uniform_i32((int)(threadIdx.x / WAVE));
    const int lane = lane_id();
    double *cextra = reinterpret_cast<double *>(smem) + (size_t)wave * M;
    T *gxs = reinterpret_cast<T *>(reinterpret_cast<double *>(smem) + (size_t)BWD_WAVES * M);  // [4][64][NXREG]
    const int64_t e_id = (int64_t)blockIdx.x * BWD_WAVES + wave;
    const bool bvalid = e_id < B * K;
    const int64_t b = bvalid ? e_id / K : 0, j = bvalid ? e_id % K : 0;
    const bool want_gx = gx_partial != nullptr || grad_X_direct != nullptr;
    Program<T> p;
    if (bvalid) p = load_program<T>(ops, ch, (const T *)nullptr, b, 1, B, 1, B, M, n_x);
    else { p.L = 0; p.w0 = p.w1 = 0; p.cmask0 = p.cmask1 = 0; p.root = 0; p.verr = 0; p.c0 = p.c1 = T(0); }
    for (int k = lane; k < M; k += WAVE) cextra[k] = 0.0;
    GradAcc<T> acc;
#pragma unroll
    for (int q = 0; q < NCREG; ++q) acc.cacc[q] = 0.0;
    int32_t err = 0;
    const int64_t n_tiles = (N + TILE - 1) / TILE;
    for (int64_t t = 0; t < n_tiles; ++t) {
        const int64_t n = t * TILE + lane;
        const bool active = bvalid && n < N;
        const int64_t nn = active ? n : 0;
#pragma unroll
        for (int q = 0; q < NXREG; ++q) acc.gxl[q] = T(0);
        for (int k = p.L; k < M; ++k) {
            T g = active ? grad_cache[(((int64_t)k * N + n) * B + b) * K + j] : T(0);
            if (g != T(0)) err = err > ERR_GRAD_NO_OP ? err : ERR_GRAD_NO_OP;
        }
        auto val = [&](int k) -> T { return cache[(((int64_t)k * N + nn) * B + b) * K + j]; };
        auto adj = [&](int k) -> T & { return grad_cache[(((int64_t)k * N + nn) * B + b) * K + j]; };
        int32_t e = backward_tile<T, true>(p, val, adj, p.L, active, acc, cextra, n_x,
                                           grad_X_direct ? grad_X_direct + nn * n_x : nullptr);
        err = e > err ? e : err;
        if (want_gx && n_x <= NXREG) {
#pragma unroll
            for (int q = 0; q < NXREG; ++q) gxs[(wave * WAVE + lane) * NXREG + q] = acc.gxl[q];
            __syncthreads();
            if (wave == 0 && n < N) {
                for (int v = 0; v < n_x; ++v) {
                    T s = gxs[(0 * WAVE + lane) * NXREG + v];
                    for (int w2 = 1; w2 < BWD_WAVES; ++w2) s += gxs[(w2 * WAVE + lane) * NXREG + v];
                    gx_partial[(int64_t)blockIdx.x * N * n_x + n * n_x + v] = s;
                }
            }
            __syncthreads();
        }
    }
    if (!bvalid) return;
    double lo, hi;
    finish_constants<T>(p, acc, cextra, lo, hi);
    if (lane < M) grad_C[((int64_t)lane * B + b) * K + j] = (T)lo;
    if (lane + WAVE < M) grad_C[((int64_t)(lane + WAVE) * B + b) * K + j] = (T)hi;
    err = wave_max_i32(err);
    if (lane == 0 && err) atomicMax(err_flag, err);
}

// Sums P partial rows of length len in FP64, in row order.
template <typename T>
__global__ void k_reduce_partials(const T *__restrict__ part, int64_t P, int64_t len, T *__restrict__ out,
                                  int accumulate) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= len) return;
    double s = accumulate ? (double)out[i] : 0.0;
    for (int64_t r = 0; r < P; ++r) s += (double)part[r * len + i];
    out[i] = (T)s;
}

// ---------------------------------------------------------------------------
// Fused root-only forward + backward over batch-first layouts.
//
// MODE 0: dy is the upstream gradient of the root output (B,N).
// MODE 1: dy is the target Y (B,N); the upstream gradient is the MSE
//         closure's (2*(yhat - Y)) * (1/N) (torch's mse_loss_backward and
//         mean backward, symbolic_torch/optimization.py:87-91) and
//         loss_out[b] = mean_n (yhat - Y)^2 accumulated in FP64.
// FWD_ONLY: forward only (y_out and, in MODE 1, loss_out).
//
// Scheduling: persistent blocks of W waves.  Phase 1 (list == nullptr)
// assigns block i the contiguous expression range [i*per_blk, ...) and the
// waves pull expressions from it through an LDS counter; expressions longer
// than LMAX are appended to a long queue instead.  Phase 2 (list != nullptr)
// is launched with a larger LMAX and pulls ids from the queue through a
// global counter.  Variable adjoints are summed per block in LDS (float
// atomics, so the order within a block is not fixed) and written as one
// partial row per block.
// ---------------------------------------------------------------------------
struct FusedArgs {
    const void *X;
    int64_t N;
    int n_x;
    const int64_t *ops;
    const int64_t *ch;
    const void *C;
    int M;
    int64_t B;
    const void *dy;
    void *y_out;
    double *loss_out;
    void *grad_C;
    void *gx_partial;     // [gridDim][N*n_x] or nullptr
    int64_t gx_row0;      // first partial row of this launch
    void *grad_X_direct;  // atomics target when the block partial does not fit LDS
    int32_t *err_b;       // (B) or nullptr
    int32_t *flags;       // [0] = validation (max code), [1] = evaluation (max code)
    const int32_t *list;
    const int32_t *list_count;
    int32_t *queue;
    int32_t *queue_count;
    int32_t *work_counter;
    int64_t per_blk;
    void *gtape;          // global tape/adjoint slots when they do not fit LDS
};

template <typename T, int LMAX, int W, int MODE, bool FWD_ONLY, bool WANT_GX, bool GT>
__global__ __launch_bounds__(W * 64) void k_fused(FusedArgs a) {
    // GT: the tape and adjoints of each wave live in a global workspace slot
    // (used when LMAX is too long for LDS); otherwise in LDS.
    extern __shared__ __align__(16) unsigned char smem[];
    const int wave = uniform_i32((int)(threadIdx.x / WAVE));
    const int lane = lane_id();
    constexpr size_t TAPE_LDS = GT ? 0 : (size_t)W * 2 * LMAX * WAVE;
    T *tape = GT ? (T *)a.gtape + ((size_t)blockIdx.x * W + wave) * 2 * LMAX * WAVE
                 : reinterpret_cast<T *>(smem) + (size_t)wave * 2 * LMAX * WAVE;
    T *adjp = tape + LMAX * WAVE;
    double *cextra = reinterpret_cast<double *>(reinterpret_cast<T *>(smem) + TAPE_LDS) + (size_t)wave * LMAX;
    T *gx_blk = reinterpret_cast<T *>(reinterpret_cast<double *>(reinterpret_cast<T *>(smem) + TAPE_LDS) +
                                      (size_t)W * LMAX);
    const T *X = (const T *)a.X;
    const T *C = (const T *)a.C;
    const T *dy = (const T *)a.dy;
    T *y_out = (T *)a.y_out;
    T *grad_C = (T *)a.grad_C;
    const bool gx_lds = WANT_GX && a.gx_partial != nullptr;
    const int64_t gx_len = a.N * a.n_x;
    if (gx_lds)
        for (int64_t i = threadIdx.x; i < gx_len; i += W * WAVE) gx_blk[i] = T(0);
    __syncthreads();

    const int64_t n_tiles = (a.N + TILE - 1) / TILE;
    const T inv_n = T(1.0) / T(a.N);
    // Phase 1: block i owns [i*per_blk, ...) and wave w takes its slots
    // w, w+W, ...  Phase 2: waves take queue slots from a global counter.
    // Either way the slot index is a wave-uniform SGPR value.
    int64_t base = 0, count = 0;
    if (a.list == nullptr) {
        base = (int64_t)blockIdx.x * a.per_blk;
        count = a.B - base < a.per_blk ? a.B - base : a.per_blk;
        if (count < 0) count = 0;
    } else {
        count = *a.list_count;
    }
    int64_t slot = wave;
    for (;;) {
        int64_t i;
        if (a.list == nullptr) {
            i = slot;
            slot += W;
        } else {
            i = wave_fetch_add(a.work_counter);
        }
        if (i >= count) break;
        const int64_t b = (a.list == nullptr) ? base + i : (int64_t)a.list[i];
        Program<T> p = load_program<T>(a.ops, a.ch, C, b, a.M, 1, a.M, 1, a.M, a.n_x);
        if (p.verr) {
            if (lane == 0) {
                atomicMax(&a.flags[0], p.verr);
                if (a.err_b) a.err_b[b] = p.verr;
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
        for (int k = lane; k < LMAX; k += WAVE) cextra[k] = 0.0;
        GradAcc<T> acc;
#pragma unroll
        for (int q = 0; q < NCREG; ++q) acc.cacc[q] = 0.0;
        double loss = 0.0;
        int32_t err = 0;
        const int L = p.L;
        for (int64_t t = 0; t < n_tiles; ++t) {
            const int64_t n = t * TILE + lane;
            const bool active = n < a.N;
            PointCtx<T> pc = load_point<T>(X, n, active, a.n_x);
            forward_tile<T>(p, pc, tape, L);
            // L == 0 only for an all-NO_OP row whose root (index 0) is NaN.
            T yh = L > 0 ? tape[p.root * WAVE + lane] : Mth<T>::qnan();
            if (active && y_out) y_out[b * a.N + n] = yh;
            T g0 = T(0);
            if (MODE == 1) {
                T target = active ? dy[b * a.N + n] : T(0);
                T d = yh - target;
                if (active) loss += (double)d * (double)d;
                g0 = (T(2.0) * d) * inv_n;
            } else {
                g0 = active ? dy[b * a.N + n] : T(0);
            }
            if (FWD_ONLY) continue;
            if (L == 0) {
                if (active && g0 != T(0)) err = ERR_GRAD_NO_OP;
                continue;
            }
            for (int k = 0; k < L; ++k) adjp[k * WAVE + lane] = T(0);
            adjp[p.root * WAVE + lane] = active ? g0 : T(0);
#pragma unroll
            for (int q = 0; q < NXREG; ++q) acc.gxl[q] = T(0);
            auto val = [&](int k) -> T { return tape[k * WAVE + lane]; };
            auto adj = [&](int k) -> T & { return adjp[k * WAVE + lane]; };
            int32_t e = backward_tile<T, WANT_GX>(p, val, adj, L, active, acc, cextra, a.n_x,
                                                  (T *)a.grad_X_direct + (active ? n : 0) * a.n_x);
            err = e > err ? e : err;
            if (WANT_GX && a.n_x <= NXREG && active) {
                if (gx_lds) {
                    for (int v = 0; v < a.n_x; ++v)
                        if (acc.gxl[v] != T(0)) atomicAdd(&gx_blk[n * a.n_x + v], acc.gxl[v]);
                } else if (a.grad_X_direct) {
                    for (int v = 0; v < a.n_x; ++v)
                        if (acc.gxl[v] != T(0)) atomicAdd(&((T *)a.grad_X_direct)[n * a.n_x + v], acc.gxl[v]);
                }
            }
        }
        if (MODE == 1 && a.loss_out) {
            double s = wave_sum(loss);
            if (lane == 0) a.loss_out[b] = s / (double)a.N;
        }
        if (!FWD_ONLY) {
            double lo, hi;
            finish_constants<T>(p, acc, cextra, lo, hi);
            if (lane < a.M) grad_C[b * a.M + lane] = (T)lo;
            if (lane + WAVE < a.M) grad_C[b * a.M + lane + WAVE] = (T)hi;
            err = wave_max_i32(err);
            if (lane == 0) {
                if (err) atomicMax(&a.flags[1], err);
                if (a.err_b) a.err_b[b] = err;
            }
        } else if (lane == 0 && a.err_b) {
            a.err_b[b] = 0;
        }
    }
    __syncthreads();
    if (gx_lds) {
        T *dst = (T *)a.gx_partial + (a.gx_row0 + blockIdx.x) * gx_len;
        for (int64_t i = threadIdx.x; i < gx_len; i += W * WAVE) dst[i] = gx_blk[i];
    }
}

inline int ret_last() { return (int)hipGetLastError(); }

constexpr int P1_LMAX = 16;
constexpr size_t GX_LDS_MAX = 24 * 1024;

template <typename T> constexpr int p1_waves() { return sizeof(T) == 4 ? 4 : 2; }

template <typename T>
size_t fused_lds_bytes(int LMAX, int W, bool gt, bool gx_lds, int64_t gx_len) {
    return (gt ? 0 : (size_t)W * 2 * LMAX * WAVE * sizeof(T)) + (size_t)W * LMAX * sizeof(double) +
           (gx_lds ? (size_t)gx_len * sizeof(T) : 0);
}
template <typename T>
bool needs_global_tape(int LMAX, int W, bool gx_lds, int64_t gx_len) {
    return fused_lds_bytes<T>(LMAX, W, false, gx_lds, gx_len) > lds_limit();
}

template <typename T, int LMAX, int W, int MODE, bool FWD, bool GX>
int launch_fused_one(FusedArgs a, int64_t grid, hipStream_t s) {
    bool gx_lds = GX && a.gx_partial != nullptr;
    bool gt = needs_global_tape<T>(LMAX, W, gx_lds, a.N * a.n_x);
    size_t lds = fused_lds_bytes<T>(LMAX, W, gt, gx_lds, a.N * a.n_x);
    if (lds > lds_limit()) return -3;
    if (gt) {
        if (a.gtape == nullptr) return -2;
        hipLaunchKernelGGL((k_fused<T, LMAX, W, MODE, FWD, GX, true>), dim3((unsigned)grid), dim3(W * 64), lds, s, a);
    } else {
        hipLaunchKernelGGL((k_fused<T, LMAX, W, MODE, FWD, GX, false>), dim3((unsigned)grid), dim3(W * 64), lds, s,
                           a);
    }
    return ret_last();
}

inline int p2_lmax(int64_t M) { return M <= 32 ? 32 : (M <= 64 ? 64 : 128); }

template <typename T, int MODE, bool FWD, bool GX>
int launch_fused_phase(FusedArgs a, bool phase2, int lmax2, int64_t grid, hipStream_t s) {
    if (!phase2) return launch_fused_one<T, P1_LMAX, p1_waves<T>(), MODE, FWD, GX>(a, grid, s);
    if (lmax2 <= 32) return launch_fused_one<T, 32, 1, MODE, FWD, GX>(a, grid, s);
    if (lmax2 <= 64) return launch_fused_one<T, 64, 1, MODE, FWD, GX>(a, grid, s);
    return launch_fused_one<T, 128, 1, MODE, FWD, GX>(a, grid, s);
}

template <typename T, int MODE, bool FWD>
int launch_fused_gx(FusedArgs a, bool want_gx, bool phase2, int lmax2, int64_t grid, hipStream_t s) {
    if (want_gx) return launch_fused_phase<T, MODE, FWD, true>(a, phase2, lmax2, grid, s);
    return launch_fused_phase<T, MODE, FWD, false>(a, phase2, lmax2, grid, s);
}

template <typename T>
int launch_fused_disp(FusedArgs a, int mode, bool fwd_only, bool want_gx, bool phase2, int lmax2, int64_t grid,
                      hipStream_t s) {
    if (mode == 1) {
        if (fwd_only) return launch_fused_gx<T, 1, true>(a, false, phase2, lmax2, grid, s);
        return launch_fused_gx<T, 1, false>(a, want_gx, phase2, lmax2, grid, s);
    }
    if (fwd_only) return launch_fused_gx<T, 0, true>(a, false, phase2, lmax2, grid, s);
    return launch_fused_gx<T, 0, false>(a, want_gx, phase2, lmax2, grid, s);
}

// Grid sizes of the two fused phases (persistent blocks; phase-2 blocks hold
// one wave each).
template <typename T>
void fused_grids(int64_t B, int64_t &g1, int64_t &g2) {
    int64_t cus = num_cus();
    const int w1 = p1_waves<T>();
    g1 = cus * (sizeof(T) == 4 ? 4 : 8);
    int64_t need = (B + w1 - 1) / w1;
    if (g1 > need) g1 = need > 0 ? need : 1;
    g2 = cus * 8;
}

// Workspace layout: [queue: B int32][4 counters] | gx partial rows | global tape.
struct FusedWs {
    size_t gx_off, gt_off, total;
    bool gx_lds, gt;
};
template <typename T>
FusedWs fused_ws_layout(int64_t N, int64_t n_x, int64_t B, int64_t M, bool want_gx) {
    int64_t g1, g2;
    fused_grids<T>(B, g1, g2);
    FusedWs w;
    const int64_t gx_len = N * n_x;
    w.gx_lds = want_gx && n_x <= NXREG && (size_t)gx_len * sizeof(T) <= GX_LDS_MAX;
    w.gt = M > P1_LMAX && needs_global_tape<T>(p2_lmax(M), 1, w.gx_lds, gx_len);
    size_t off = (((size_t)B * 4 + 16) + 255) / 256 * 256;
    w.gx_off = off;
    if (w.gx_lds) off += (size_t)(g1 + g2) * gx_len * sizeof(T);
    off = (off + 255) / 256 * 256;
    w.gt_off = off;
    if (w.gt) off += (size_t)g2 * 2 * p2_lmax(M) * WAVE * sizeof(T);
    w.total = off + 256;
    return w;
}

template <typename T>
int fused_root(const T *X, int64_t N, int64_t n_x, const int64_t *ops, const int64_t *ch, const T *C, int64_t M,
               int64_t B, const T *dy, int mode, int fwd_only, T *y_out, double *loss_out, T *grad_C, T *grad_X,
               int32_t *err_b, int32_t *flags, void *workspace, int64_t workspace_bytes, hipStream_t s) {
    if (M <= 0 || M > 128 || B < 0 || N <= 0 || n_x < 0) return -1;
    if (B == 0) return 0;
    int64_t g1, g2;
    fused_grids<T>(B, g1, g2);
    const int64_t gx_len = N * n_x;
    const bool want_gx = grad_X != nullptr && !fwd_only;
    FusedWs L = fused_ws_layout<T>(N, n_x, B, M, want_gx);
    if (workspace_bytes < (int64_t)L.total) return -2;
    char *ws = (char *)workspace;
    int32_t *queue = (int32_t *)ws;
    int32_t *counters = (int32_t *)(ws + (size_t)B * 4);
    T *gx_part = L.gx_lds ? (T *)(ws + L.gx_off) : nullptr;

    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, counters, 0, (int64_t)4);
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, flags, 0, (int64_t)2);
    if (want_gx && !L.gx_lds) {
        // direct atomics into grad_X
        int64_t nb = (gx_len + 255) / 256;
        if (sizeof(T) == 4)
            hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)nb), dim3(256), 0, s, (int32_t *)grad_X, 0, gx_len);
        else
            hipLaunchKernelGGL(k_fill_u64, dim3((unsigned)nb), dim3(256), 0, s, (uint64_t *)grad_X, 0ull, gx_len);
    }
    FusedArgs a;
    memset(&a, 0, sizeof(a));
    a.X = X; a.N = N; a.n_x = (int)n_x; a.ops = ops; a.ch = ch; a.C = C; a.M = (int)M; a.B = B; a.dy = dy;
    a.y_out = y_out; a.loss_out = loss_out; a.grad_C = grad_C;
    a.gx_partial = gx_part;
    a.gx_row0 = 0;
    a.grad_X_direct = (want_gx && !L.gx_lds) ? grad_X : nullptr;
    a.err_b = err_b; a.flags = flags;
    a.list = nullptr; a.list_count = nullptr;
    a.queue = queue; a.queue_count = counters;
    a.work_counter = counters + 1;
    a.per_blk = (B + g1 - 1) / g1;
    a.gtape = L.gt ? (void *)(ws + L.gt_off) : nullptr;
    int r = launch_fused_disp<T>(a, mode, fwd_only != 0, want_gx, false, 0, g1, s);
    if (r) return r;
    if (M > P1_LMAX) {
        FusedArgs a2 = a;
        a2.list = queue;
        a2.list_count = counters;
        a2.gx_row0 = g1;
        r = launch_fused_disp<T>(a2, mode, fwd_only != 0, want_gx, true, p2_lmax(M), g2, s);
        if (r) return r;
    } else if (L.gx_lds) {
        // phase-2 rows stay unused: zero them so the reduction can read all rows
        int64_t cnt = g2 * gx_len;
        if (sizeof(T) == 4)
            hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, s,
                               (int32_t *)(gx_part + g1 * gx_len), 0, cnt);
        else
            hipLaunchKernelGGL(k_fill_u64, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, s,
                               (uint64_t *)(gx_part + g1 * gx_len), 0ull, cnt);
    }
    if (L.gx_lds) {
        hipLaunchKernelGGL(k_reduce_partials<T>, dim3((unsigned)((gx_len + 255) / 256)), dim3(256), 0, s,
                           (const T *)gx_part, g1 + g2, gx_len, grad_X, 0);
    }
    return ret_last();
}

template <typename T>
int64_t fused_ws_bytes(int64_t N, int64_t n_x, int64_t B, int64_t M, int want_gx) {
    return (int64_t)fused_ws_layout<T>(N, n_x, B, M, want_gx != 0).total;
}

template <typename T>
int forward_mb(const T *X, int64_t N, int64_t n_x, const int64_t *ops, const int64_t *ch, const T *C, int64_t M,
               int64_t B, int64_t K, T *cache, hipStream_t s) {
    if (M <= 0 || M > 128 || N < 0 || B < 0 || n_x < 0 || K <= 0) return -1;
    if (N == 0 || B == 0) return 0;
    int64_t n_tiles = (N + TILE - 1) / TILE;
    int64_t nblk = n_tiles * B * K;
    if (nblk > 0x7fffffffLL) return -1;
    size_t lds = (size_t)M * WAVE * sizeof(T);
    if (lds > lds_limit()) return -3;
    hipLaunchKernelGGL(k_full_forward<T>, dim3((unsigned)nblk), dim3(64), lds, s, X, N, (int)n_x, ops, ch, C,
                       (int)M, B, K, n_tiles, cache);
    return ret_last();
}

template <typename T>
int64_t backward_mb_ws_bytes(int64_t B, int64_t N, int64_t n_x, int64_t K = 1) {
    int64_t nblk = (B * K + BWD_WAVES - 1) / BWD_WAVES;
    return nblk * N * n_x * (int64_t)sizeof(T) + 256;
}

template <typename T>
int backward_mb(T *grad_out, const T *cache, const int64_t *ops, const int64_t *ch, int64_t M, int64_t B,
                int64_t K, int64_t N, int64_t n_x, T *grad_C, T *grad_X, void *workspace, int64_t ws_bytes,
                int32_t *d_err, hipStream_t s) {
    if (M <= 0 || M > 128 || N < 0 || B < 0 || n_x < 0 || K <= 0) return -1;
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, d_err, 0, (int64_t)1);
    int64_t nblk = (B * K + BWD_WAVES - 1) / BWD_WAVES;
    const bool gx_regs = n_x <= NXREG;
    if (grad_X) {
        int64_t len = N * n_x;
        if (len > 0) {
            if (sizeof(T) == 4)
                hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, s, (int32_t *)grad_X,
                                   0, len);
            else
                hipLaunchKernelGGL(k_fill_u64, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, s,
                                   (uint64_t *)grad_X, 0ull, len);
        }
    }
    if (B == 0 || N == 0) {
        int64_t len = M * B * K;
        if (len > 0) {
            if (sizeof(T) == 4)
                hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, s, (int32_t *)grad_C,
                                   0, len);
            else
                hipLaunchKernelGGL(k_fill_u64, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, s,
                                   (uint64_t *)grad_C, 0ull, len);
        }
        return ret_last();
    }
    T *gx_partial = nullptr;
    if (grad_X && gx_regs) {
        if (ws_bytes < backward_mb_ws_bytes<T>(B, N, n_x, K)) return -2;
        gx_partial = (T *)workspace;
    }
    size_t lds = (size_t)BWD_WAVES * M * sizeof(double) + (size_t)BWD_WAVES * WAVE * NXREG * sizeof(T);
    if (lds > lds_limit()) return -3;
    hipLaunchKernelGGL(k_full_backward<T>, dim3((unsigned)nblk), dim3(BWD_WAVES * 64), lds, s, grad_out, cache, ops,
                       ch, (int)M, B, K, N, (int)n_x, grad_C, gx_partial, (grad_X && !gx_regs) ? grad_X : nullptr,
                       d_err);
    if (gx_partial) {
        int64_t len = N * n_x;
        hipLaunchKernelGGL(k_reduce_partials<T>, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, s,
                           (const T *)gx_partial, nblk, len, grad_X, 0);
    }
    return ret_last();
}

}  // namespace

extern "C" {

int sr_validate_mb(const int64_t *ops, const int64_t *ch, int64_t M, int64_t B, unsigned long long *d_first,
                   void *stream) {
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, s, (uint64_t *)d_first, ~0ull, (int64_t)1);
    int64_t tot = M * B;
    if (tot > 0)
        hipLaunchKernelGGL(k_validate_mb, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, ops, ch, M, B, d_first);
    return ret_last();
}

int sr_forward_mb_f32(const float *X, int64_t N, int64_t n_x, const int64_t *ops, const int64_t *ch, const float *C,
                      int64_t M, int64_t B, float *cache, void *stream) {
    return forward_mb<float>(X, N, n_x, ops, ch, C, M, B, 1, cache, (hipStream_t)stream);
}
int sr_forward_mb_f64(const double *X, int64_t N, int64_t n_x, const int64_t *ops, const int64_t *ch,
                      const double *C, int64_t M, int64_t B, double *cache, void *stream) {
    return forward_mb<double>(X, N, n_x, ops, ch, C, M, B, 1, cache, (hipStream_t)stream);
}

int64_t sr_backward_mb_workspace_bytes(int64_t B, int64_t N, int64_t n_x, int is_f64) {
    return is_f64 ? backward_mb_ws_bytes<double>(B, N, n_x) : backward_mb_ws_bytes<float>(B, N, n_x);
}

int sr_backward_mb_f32(float *grad_out, const float *cache, const int64_t *ops, const int64_t *ch, int64_t M,
                       int64_t B, int64_t N, int64_t n_x, float *grad_C, float *grad_X, void *workspace,
                       int64_t ws_bytes, int32_t *d_err, void *stream) {
    return backward_mb<float>(grad_out, cache, ops, ch, M, B, 1, N, n_x, grad_C, grad_X, workspace, ws_bytes, d_err,
                              (hipStream_t)stream);
}
int sr_backward_mb_f64(double *grad_out, const double *cache, const int64_t *ops, const int64_t *ch, int64_t M,
                       int64_t B, int64_t N, int64_t n_x, double *grad_C, double *grad_X, void *workspace,
                       int64_t ws_bytes, int32_t *d_err, void *stream) {
    return backward_mb<double>(grad_out, cache, ops, ch, M, B, 1, N, n_x, grad_C, grad_X, workspace, ws_bytes, d_err,
                               (hipStream_t)stream);
}

int sr_multi_forward_mb_f32(const float *X, int64_t N, int64_t n_x, const int64_t *ops, const int64_t *ch,
                            const float *C, int64_t M, int64_t B, int64_t K, float *cache, void *stream) {
    return forward_mb<float>(X, N, n_x, ops, ch, C, M, B, K, cache, (hipStream_t)stream);
}
int sr_multi_forward_mb_f64(const double *X, int64_t N, int64_t n_x, const int64_t *ops, const int64_t *ch,
                            const double *C, int64_t M, int64_t B, int64_t K, double *cache, void *stream) {
    return forward_mb<double>(X, N, n_x, ops, ch, C, M, B, K, cache, (hipStream_t)stream);
}
int64_t sr_multi_backward_mb_workspace_bytes(int64_t B, int64_t N, int64_t n_x, int64_t K, int is_f64) {
    return is_f64 ? backward_mb_ws_bytes<double>(B, N, n_x, K) : backward_mb_ws_bytes<float>(B, N, n_x, K);
}
int sr_multi_backward_mb_f32(float *grad_out, const float *cache, const int64_t *ops, const int64_t *ch,
                             int64_t M, int64_t B, int64_t N, int64_t n_x, int64_t K, float *grad_C, float *grad_X,
                             void *workspace, int64_t ws_bytes, int32_t *d_err, void *stream) {
    return backward_mb<float>(grad_out, cache, ops, ch, M, B, K, N, n_x, grad_C, grad_X, workspace, ws_bytes, d_err,
                              (hipStream_t)stream);
}
int sr_multi_backward_mb_f64(double *grad_out, const double *cache, const int64_t *ops, const int64_t *ch,
                             int64_t M, int64_t B, int64_t N, int64_t n_x, int64_t K, double *grad_C, double *grad_X,
                             void *workspace, int64_t ws_bytes, int32_t *d_err, void *stream) {
    return backward_mb<double>(grad_out, cache, ops, ch, M, B, K, N, n_x, grad_C, grad_X, workspace, ws_bytes, d_err,
                               (hipStream_t)stream);
}

int64_t sr_fused_workspace_bytes(int64_t N, int64_t n_x, int64_t B, int64_t M, int want_gx, int is_f64) {
    return is_f64 ? fused_ws_bytes<double>(N, n_x, B, M, want_gx) : fused_ws_bytes<float>(N, n_x, B, M, want_gx);
}

int sr_fused_root_f32(const float *X, int64_t N, int64_t n_x, const int64_t *ops, const int64_t *ch, const float *C,
                      int64_t M, int64_t B, const float *dy, int mode, int fwd_only, float *y_out, double *loss_out,
                      float *grad_C, float *grad_X, int32_t *err_b, int32_t *flags, void *workspace,
                      int64_t workspace_bytes, void *stream) {
    return fused_root<float>(X, N, n_x, ops, ch, C, M, B, dy, mode, fwd_only, y_out, loss_out, grad_C, grad_X, err_b,
                             flags, workspace, workspace_bytes, (hipStream_t)stream);
}
int sr_fused_root_f64(const double *X, int64_t N, int64_t n_x, const int64_t *ops, const int64_t *ch,
                      const double *C, int64_t M, int64_t B, const double *dy, int mode, int fwd_only,
                      double *y_out, double *loss_out, double *grad_C, double *grad_X, int32_t *err_b,
                      int32_t *flags, void *workspace, int64_t workspace_bytes, void *stream) {
    return fused_root<double>(X, N, n_x, ops, ch, C, M, B, dy, mode, fwd_only, y_out, loss_out, grad_C, grad_X,
                              err_b, flags, workspace, workspace_bytes, (hipStream_t)stream);
}

}  // extern "C"
