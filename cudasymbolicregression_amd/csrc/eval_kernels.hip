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
//
// (The fused root-only forward + backward is in fused_kernels.hip.)
//
// Layout conventions: these backend ("mb") entry points take the reference's
// node-first tensors Ops (M,B), Ch (M,B,2), C (M,B), cache (M,N,B).
#include "sr_wave.h"
#include "sr_device_info.h"
#include "sr_streams.h"
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
// Row-buffer forward (with a workspace).  Most of the cache is NO_OP padding,
// and the grouped kernel's stores cover EW * 4 contiguous bytes per point,
// which holds HBM writes to about a sixth of their peak.  With a workspace the
// forward runs in two passes over groups of FG = 64 consecutive expressions:
//   pass 1 (k_fwd_rows) evaluates every expression with its rows k < L_e in
//     the row buffer as the tape (N contiguous values per row);
//   pass 2 (k_fwd_transpose) writes the cache one (k, n) row of 64
//     expressions at a time, 64 * sizeof(T) contiguous bytes, NaN for k >= L_e.
// k_fwd_plan / k_fwd_scan place the groups in the row buffer.  A group whose
// rows do not fit in the buffer is evaluated with the cache itself as the
// tape and skips pass 2, so any workspace size gives the same cache.
// Workspace: goff int64[G + 1] | eoff int64[E] | Lw uint8[E] | rows T[cap].
// ---------------------------------------------------------------------------
constexpr int FG = WAVE;  // expressions per row-buffer group (one per lane)
constexpr int64_t FWD_ROWS_SHARE = 4;  // recommended row buffer: 1 / FWD_ROWS_SHARE of the cache

struct FwdRows {
    const int64_t *goff;  // group g's rows start at goff[g] and end at goff[g + 1]
    const int64_t *eoff;  // expression e's rows start at goff[e / FG] + eoff[e]
    const uint8_t *Lw;    // rows of expression e (its highest non-NO_OP node + 1)
    void *rows;
    int64_t cap;          // elements of rows; -1: no row buffer
};

inline int64_t align256(int64_t x) { return (x + 255) / 256 * 256; }
inline int64_t fwd_index_bytes(int64_t E) {
    const int64_t G = (E + FG - 1) / FG;
    return align256(8 * (G + 1)) + align256(8 * E) + align256(E);
}

// One wave per group: L of every expression, offsets inside the group, and
// the group's size in goff[g] (made exclusive offsets by k_fwd_scan).
__global__ __launch_bounds__(64) void k_fwd_plan(const int64_t *__restrict__ ops, int M, int64_t B, int64_t K,
                                                 int64_t N, int64_t *__restrict__ goff, int64_t *__restrict__ eoff,
                                                 uint8_t *__restrict__ Lw) {
    const int lane = lane_id();
    const int64_t E = B * K;
    const int64_t e = (int64_t)blockIdx.x * FG + lane;
    int L = 0;
    if (e < E) {
        const int64_t b = e / K;
        for (int k = 0; k < M; ++k) L = ops[(int64_t)k * B + b] != NO_OP ? k + 1 : L;
    }
    const int64_t size = (int64_t)L * N;
    int64_t incl = size;
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) {
        const int64_t up = __shfl_up(incl, d, WAVE);
        if (lane >= d) incl += up;
    }
    if (e < E) {
        eoff[e] = incl - size;
        Lw[e] = (uint8_t)L;
    }
    if (lane == WAVE - 1) goff[blockIdx.x] = incl;
}

// One block: goff[0..G) sizes -> exclusive offsets, goff[G] = total.
__global__ __launch_bounds__(1024) void k_fwd_scan(int64_t *__restrict__ goff, int64_t G) {
    __shared__ int64_t wsum[16];
    __shared__ int64_t carry_s;
    const int t = (int)threadIdx.x, lane = lane_id(), w = t / WAVE;
    if (t == 0) carry_s = 0;
    __syncthreads();
    for (int64_t base = 0; base < G; base += 1024) {
        const int64_t i = base + t;
        const int64_t v = i < G ? goff[i] : 0;
        int64_t incl = v;
#pragma unroll
        for (int d = 1; d < WAVE; d <<= 1) {
            const int64_t up = __shfl_up(incl, d, WAVE);
            if (lane >= d) incl += up;
        }
        if (lane == WAVE - 1) wsum[w] = incl;
        __syncthreads();
        int64_t before = carry_s;
        for (int q = 0; q < w; ++q) before += wsum[q];
        if (i < G) goff[i] = before + incl - v;
        __syncthreads();
        if (t == 1023) carry_s = before + incl;
        __syncthreads();
    }
    if (t == 0) goff[G] = carry_s;
}

// ---------------------------------------------------------------------------
// Full-cache forward with coalesced cache writes.  The cache (M,N,B[,K]) has
// the expressions innermost, so one wave writing one expression's points
// touches a different cache line per lane.  Here a block of EW waves
// evaluates EW consecutive expressions e (= b*K + j) over one 64-point tile,
// each into its own LDS tape, and then the whole block writes node row by
// node row with the expression index fastest: every group of EW threads
// stores EW contiguous values.  Blocks are XCD-remapped so that consecutive
// expression groups of one tile run on the same XCD and fill whole L2 lines
// together.  Tapes are EW apart by M*64 + 1 elements (conflict-free
// transposed reads).
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(512) void k_full_forward_grp(const T *__restrict__ X, int64_t N, int n_x,
                                                         const int64_t *__restrict__ ops,
                                                         const int64_t *__restrict__ ch, const T *__restrict__ C,
                                                         int M, int64_t B, int64_t K, int64_t n_tiles, int EW,
                                                         T *__restrict__ cache) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ int Ls[8];
    const int wave = uniform_i32((int)(threadIdx.x / WAVE));
    const int lane = lane_id();
    const int64_t E = B * K;
    const int64_t groups = (E + EW - 1) / EW;
    const int64_t id = xcd_remap(blockIdx.x, (int64_t)gridDim.x);
    const int64_t tile = id / groups, g = id % groups;
    if (tile >= n_tiles) return;  // whole block
    const int64_t e0 = g * EW;
    const size_t tstride = (size_t)M * WAVE + 1;
    T *tape = reinterpret_cast<T *>(smem) + (size_t)wave * tstride;
    {
        const int64_t e = e0 + wave;
        int L = 0;
        if (e < E) {
            const int64_t b = e / K, j = e % K;
            Program<T> p = load_program<T>(ops, ch, C + j, b, 1, B, K, B * K, M, n_x);
            const int64_t n = tile * TILE + lane;
            PointCtx<T> pc = load_point<T>(X, n, n < N, n_x);
            forward_tile<T>(p, pc, tape, p.L);
            L = p.L;
        }
        if (lane == 0) Ls[wave] = L;
    }
    __syncthreads();
    const int t = (int)threadIdx.x;
    const int el = t % EW, nl = t / EW;  // nl < 64 since blockDim = 64 * EW
    const int64_t e = e0 + el, n = tile * TILE + nl;
    if (e >= E || n >= N) return;
    const T *src = reinterpret_cast<const T *>(smem) + (size_t)el * tstride + nl;
    const int L = Ls[el];
    const T nan = Mth<T>::qnan();
    for (int k = 0; k < M; ++k) cache[((int64_t)k * N + n) * E + e] = k < L ? src[(size_t)k * WAVE] : nan;
}

// Pass 1: one wave per (expression e_begin + i, i < e_count, and FR_TILES
// consecutive 64-point tiles), FR_WAVES per block.  A program of at most FR_LROWS nodes runs on an
// LDS tape and then writes its rows to the row buffer; a longer one uses its
// rows in the row buffer as the tape (lanes = points), so the LDS per wave
// stays small and the pass runs at full occupancy.  A group with no room in the
// buffer is evaluated with the cache itself as the tape (node k of expression
// e at point n in its final place, then the NaN padding) and skips pass 2.
constexpr int FR_WAVES = 4;
constexpr int FR_LROWS = 16;  // programs of at most this many nodes use an LDS tape
constexpr int FR_TILES = 4;   // consecutive point tiles per wave (one program load)
template <typename T>
__global__ __launch_bounds__(FR_WAVES * 64) void k_fwd_rows(const T *__restrict__ X, int64_t N, int n_x,
                                                           const int64_t *__restrict__ ops,
                                                           const int64_t *__restrict__ ch, const T *__restrict__ C,
                                                           int M, int64_t B, int64_t K, int64_t n_tiles, int64_t e_begin,
                                                           int64_t e_count, FwdRows fr, T *__restrict__ cache) {
    __shared__ T ltape[FR_WAVES][FR_LROWS * WAVE];
    const int wave = uniform_i32((int)(threadIdx.x / WAVE));
    const int lane = lane_id();
    const int64_t E = B * K;
    const int64_t id = (int64_t)blockIdx.x * FR_WAVES + wave;
    const int64_t e = e_begin + id % e_count, tile0 = id / e_count * FR_TILES;
    if (tile0 >= n_tiles) return;  // whole wave
    const int64_t b = e / K, j = e % K, grp = e / FG;
    const Program<T> p = load_program<T>(ops, ch, C + j, b, 1, B, K, B * K, M, n_x);
    const bool fits = fr.goff[grp + 1] <= fr.cap;
    const int L = fits ? (p.L < (int)fr.Lw[e] ? p.L : (int)fr.Lw[e]) : p.L;
    // the program is loaded once for FR_TILES consecutive tiles
    for (int64_t tile = tile0; tile < tile0 + FR_TILES && tile < n_tiles; ++tile) {
        const int64_t n = tile * TILE + lane;
        if (n >= N) break;  // no wave-wide operation follows
        const PointCtx<T> pc = load_point<T>(X, n, true, n_x);
        // Rows are clamped below L: a valid program only names children
        // below their parent, and an unvalidated one stays in bounds.
        if (fits) {
            T *t = reinterpret_cast<T *>(fr.rows) + fr.goff[grp] + fr.eoff[e] + n;
            if (L <= FR_LROWS) {  // short program: LDS tape, then its rows out
                T *lt = ltape[wave];
                forward_tile_at<T>(p, pc, [&](int k) -> T & { return lt[(k < L ? k : L - 1) * WAVE + lane]; }, L);
                for (int k = 0; k < L; ++k) t[(int64_t)k * N] = lt[k * WAVE + lane];
            } else {
                forward_tile_at<T>(p, pc, [&](int k) -> T & { return t[(int64_t)(k < L ? k : L - 1) * N]; }, L);
            }
        } else if (cache != nullptr) {  // (no cache: the rows only)
            T *t = cache + n * E + e;
            const int64_t rs = N * E;
            forward_tile_at<T>(p, pc, [&](int k) -> T & { return t[(int64_t)(k < L ? k : L - 1) * rs]; }, L);
            const T nan = Mth<T>::qnan();
            for (int k = L; k < M; ++k) t[(int64_t)k * rs] = nan;
        }
    }
}

// Pass 2: block (group g_begin + i, i < g_count, and 64-point tile),
// TR_WAVES waves.  Per node row k:
// the group's 64 rows (lanes = points) into an LDS tile, then the cache rows
// (k, n) with lanes = expressions, NaN for k >= L_e.  The loads of row k + 1
// are issued before row k's stores, so they are in flight while those drain.
// Rows past every L_e are NaN without the tile.  The tile is 64 x 65 (odd
// stride: the transposed reads are conflict-free).
constexpr int TR_WAVES = 8;
template <typename T>
__global__ __launch_bounds__(TR_WAVES * 64) void k_fwd_transpose(int M, int64_t N, int64_t E, int64_t g_begin,
                                                                int64_t g_count, FwdRows fr, T *__restrict__ cache) {
    __shared__ T tile[FG][WAVE + 1];
    __shared__ int64_t base_s[FG];
    __shared__ int L_s[FG];
    const int wave = uniform_i32((int)(threadIdx.x / WAVE));
    const int lane = lane_id();
    const int64_t g = g_begin + (int64_t)blockIdx.x % g_count, tl = (int64_t)blockIdx.x / g_count;
    if (fr.goff[g + 1] > fr.cap) return;  // written by pass 1
    const int64_t e = g * FG + lane, n0 = tl * WAVE;
    if (wave == 0) {
        L_s[lane] = e < E ? (int)fr.Lw[e] : 0;
        base_s[lane] = e < E ? fr.goff[g] + fr.eoff[e] : 0;
    }
    __syncthreads();
    int Lmax = 0;
    for (int q = 0; q < FG; ++q) Lmax = L_s[q] > Lmax ? L_s[q] : Lmax;
    const T *rows = reinterpret_cast<const T *>(fr.rows);
    const int64_t gbase = fr.goff[g];  // a valid element whenever Lmax > 0
    const T nan = Mth<T>::qnan();
    const int myL = L_s[lane];
    const int64_t n = n0 + lane;
    constexpr int QW = FG / TR_WAVES;
    T v[QW];
    // expression q's row k over the tile's points (every load at an
    // in-bounds address of the group; unused values are dropped)
    auto load_row = [&](int k) {
#pragma unroll
        for (int i = 0; i < QW; ++i) {
            const int q = wave + i * TR_WAVES;
            const bool ok = k < L_s[q] && n < N;
            v[i] = rows[ok ? base_s[q] + (int64_t)k * N + n : gbase];
        }
    };
    if (Lmax > 0) load_row(0);
    for (int k = 0; k < M; ++k) {
        if (k < Lmax) {
#pragma unroll
            for (int i = 0; i < QW; ++i) {
                const int q = wave + i * TR_WAVES;
                if (k < L_s[q] && n < N) tile[q][lane] = v[i];
            }
            __syncthreads();
            if (k + 1 < Lmax) load_row(k + 1);
            for (int pp = wave; pp < WAVE; pp += TR_WAVES) {
                const int64_t nn = n0 + pp;
                if (nn < N && e < E) cache[((int64_t)k * N + nn) * E + e] = k < myL ? tile[lane][pp] : nan;
            }
            __syncthreads();
        } else {
            for (int pp = wave; pp < WAVE; pp += TR_WAVES) {
                const int64_t nn = n0 + pp;
                if (nn < N && e < E) cache[((int64_t)k * N + nn) * E + e] = nan;
            }
        }
    }
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

// The block's per-tile grad_X partial: each wave's variable adjoints at its
// lane's point (gxl, from the sweep when n_x <= NXREG, else recomputed per
// chunk of NXREG variables with var_adjoints) go through LDS (gxs), and wave
// 0 adds the BWD_WAVES waves in wave order into the block's partial row
// gx_row (N x n_x).  Every wave of the block calls it (it has barriers); the
// caller puts a barrier between it and the next write of gxs.  MANYX
// (n_x > NXREG) is a separate instantiation of the kernels, so the common
// case carries no code (or registers) for the chunked sums.
template <typename T>
__device__ inline void flush_gx_chunk(const T (&gxl)[NXREG], T *gxs, int wave, int lane, int64_t n, int64_t N,
                                      int n_x, int vb, int vn, T *gx_row) {
#pragma unroll
    for (int q = 0; q < NXREG; ++q) gxs[(wave * WAVE + lane) * NXREG + q] = gxl[q];
    __syncthreads();
    if (wave == 0 && n < N) {
        for (int v = 0; v < vn; ++v) {
            T sum = gxs[(0 * WAVE + lane) * NXREG + v];
            for (int w2 = 1; w2 < BWD_WAVES; ++w2) sum += gxs[(w2 * WAVE + lane) * NXREG + v];
            gx_row[n * n_x + vb + v] = sum;
        }
    }
}
template <typename T, bool MANYX, typename Adj>
__device__ inline void flush_gx(const Program<T> &p, Adj adj, int L, bool active, int n_x, T (&gxl)[NXREG], T *gxs,
                                int wave, int lane, int64_t n, int64_t N, T *gx_row) {
    if constexpr (!MANYX) {
        flush_gx_chunk<T>(gxl, gxs, wave, lane, n, N, n_x, 0, n_x, gx_row);
    } else {
        for (int vb = 0; vb < n_x; vb += NXREG) {
            if (vb > 0) __syncthreads();  // gxs is rewritten
            var_adjoints<T>(p, adj, L, active, vb, gxl);
            flush_gx_chunk<T>(gxl, gxs, wave, lane, n, N, n_x, vb, n_x - vb < NXREG ? n_x - vb : NXREG, gx_row);
        }
    }
}

template <typename T, bool MANYX>
__global__ __launch_bounds__(BWD_WAVES * 64) void k_full_backward(
    T *__restrict__ grad_cache, const T *__restrict__ cache, const int64_t *__restrict__ ops,
    const int64_t *__restrict__ ch, int M, int64_t B, int64_t K, int64_t N, int n_x, T *__restrict__ grad_C,
    T *__restrict__ gx_partial, int32_t *err_flag) {
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
    const bool want_gx = gx_partial != nullptr;
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
        int32_t e = backward_tile<T, true>(p, val, adj, p.L, active, acc, cextra, n_x);
        err = e > err ? e : err;
        if (want_gx) {
            flush_gx<T, MANYX>(p, adj, p.L, active, n_x, acc.gxl, gxs, wave, lane, n, N,
                               gx_partial + (int64_t)blockIdx.x * N * n_x);
            __syncthreads();  // gxs is rewritten for the next tile
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

// Full-cache backward with coalesced reads: the block's BWD_WAVES
// consecutive expressions load each 64-point tile of their upstream
// gradients (all M rows; rows past an expression's length must be zero, else
// error 201) and of their cached values (rows < L) into LDS with the
// expression index fastest, then every wave runs its reverse sweep on its
// LDS tapes.  The adjoints stay in LDS: grad_cache is the caller's scratch
// copy of the upstream gradient and is only read.  Outputs and reductions
// as in k_full_backward.
constexpr int BWD_BATCH = 16;
constexpr int64_t BWD_CHUNKS = 1;  // pipeline chunks of the gathered backward
constexpr int BWR_WPE = 8;  // waves per SIMD the FP32 row-tape sweep's registers are sized for (FP64 would spill)
constexpr int BWR_U = 2;    // point tiles the row-tape sweep runs at once (node by node)

// Reverse sweep of one expression over BWR_U point tiles at once, on its
// gathered rows: vr (values) and gr (adjoints, accumulated in place), row k
// of a tile at k * N + n.  Node by node, each step touches the BWR_U tiles'
// rows together, so their loads are in flight at once.  For each tile the
// steps are backward_tile's, in the same node order, and a constant's FP64
// sum takes its tiles in tile order, so every result equals tile-by-tile
// sweeps (as values: a NaN's payload may differ, the compiler being free to
// order the operands of an addition).  Variable adjoints are not summed here (var_adjoints
// reads them back from the rows).  Rows are clamped below L (a valid program
// only names children below their parent; an unvalidated one stays in
// bounds).  Returns this lane's error code.
template <typename T, int U>
__device__ inline int32_t backward_rows_batch(const Program<T> &p, const T *__restrict__ vr, T *__restrict__ gr,
                                              int64_t N, int L, const int64_t (&nu)[U], const bool (&act)[U],
                                              GradAcc<T> &acc, double *cextra) {
    const int lane = lane_id();
    int32_t err = 0;
    auto row = [&](int k) -> int64_t { return (int64_t)(k < L ? k : L - 1) * N; };
    for (int k = L - 1; k >= 0; --k) {
        const uint32_t w = p.word(k);
        const int op = node_op(w);
        if (op == UNKNOWN_LEAF || (op < SIN && op != NO_OP && op != LEARNABLE_CONSTANT) || op >= VAR_START_ID)
            continue;  // hard constants, 0-valued leaves, variables: nothing to do
        const int64_t rk = row(k);
        T g[U];
#pragma unroll
        for (int u = 0; u < U; ++u) g[u] = gr[rk + nu[u]];
        if (op == NO_OP) {
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (act[u] && g[u] != T(0)) err = err > ERR_GRAD_NO_OP ? err : ERR_GRAD_NO_OP;
            continue;
        }
        if (op == LEARNABLE_CONSTANT) {
            const int j = p.const_ordinal(k);
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const double gv = (act[u] && g[u] != T(0)) ? (double)g[u] : 0.0;
                if (j < NCREG) {
#pragma unroll
                    for (int q = 0; q < NCREG; ++q)
                        if (q == j) acc.cacc[q] += gv;
                } else {
                    const double sw = wave_sum(gv);
                    if (lane == 0) cextra[k] += sw;
                }
            }
            continue;
        }
        const int c0 = node_c0(w), c1 = node_c1(w);
        const int ar = arity_of(op);
        const int64_t r0 = row(c0), r1 = row(c1);
        T a0[U], a1[U], o0[U], o1[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            a0[u] = vr[r0 + nu[u]];
            a1[u] = ar == 2 ? vr[r1 + nu[u]] : T(0);
            o0[u] = gr[r0 + nu[u]];
            o1[u] = ar == 2 ? gr[r1 + nu[u]] : T(0);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            T go0 = T(0), go1 = T(0);
            bool add0, add1;
            const int32_t e = bwd_node<T>(op, g[u], a0[u], a1[u], go0, add0, go1, add1);
            if (act[u] && e > err) err = e;
            add0 = act[u] && add0;
            add1 = act[u] && add1 && ar == 2;
            if (ar == 2 && r0 == r1) {
                // one row takes both adds, first operand first
                T v = o0[u];
                if (add0) v += go0;
                if (add1) v += go1;
                if (add0 || add1) gr[r0 + nu[u]] = v;
            } else {
                if (add0) gr[r0 + nu[u]] = o0[u] + go0;
                if (add1) gr[r1 + nu[u]] = o1[u] + go1;
            }
        }
    }
    return err;
}

// Backward gather (with a row-buffer workspace).  The grouped backward
// below reads the upstream gradient and the cached values per (node, point)
// with the expressions of its block innermost: BWD_WAVES * sizeof(T)
// contiguous bytes.  Here one block per (group g of 64 expressions, 64-point
// tile) reads every (k, n) row of the group 64 expressions at a time, keeps
// rows k < L_e in LDS tiles, and writes them to per-expression gradient rows
// (fr.rows) and value rows (vrows), N contiguous values per row, which the
// grouped backward then reads with lanes = points.  An upstream gradient on a
// NO_OP padding row (k >= L_e) that is not zero is reported as
// ERR_GRAD_NO_OP, as the grouped backward does for the groups it reads itself
// (those without room in the buffer).
// VALS = false: only the gradient is gathered (the reverse sweep recomputes
// the forward values), so the cache is not read and vrows not written.
constexpr int GA_WAVES = 8;
template <typename T, bool VALS>
__global__ __launch_bounds__(GA_WAVES * 64) void k_bwd_gather(const T *__restrict__ grad_cache,
                                                             const T *__restrict__ cache, int M, int64_t N, int64_t E,
                                                             int64_t g_begin, int64_t g_count, FwdRows fr,
                                                             T *__restrict__ vrows, int32_t *err_flag) {
    __shared__ T gtile[FG][WAVE + 1];
    __shared__ T vtile[VALS ? FG : 1][VALS ? WAVE + 1 : 1];
    __shared__ int64_t base_s[FG];
    __shared__ int L_s[FG];
    const int wave = uniform_i32((int)(threadIdx.x / WAVE));
    const int lane = lane_id();
    const int64_t g = g_begin + (int64_t)blockIdx.x % g_count, tl = (int64_t)blockIdx.x / g_count;
    if (fr.goff[g + 1] > fr.cap) return;  // read by the grouped backward itself
    const int64_t e = g * FG + lane, n0 = tl * WAVE;
    if (wave == 0) {
        L_s[lane] = e < E ? (int)fr.Lw[e] : 0;
        base_s[lane] = e < E ? fr.goff[g] + fr.eoff[e] : 0;
    }
    __syncthreads();
    int Lmax = 0;
    for (int q = 0; q < FG; ++q) Lmax = L_s[q] > Lmax ? L_s[q] : Lmax;
    const int myL = L_s[lane];
    T *grows = reinterpret_cast<T *>(fr.rows);
    constexpr int PW = WAVE / GA_WAVES;  // points per wave
    int32_t lerr = 0;
    for (int k = 0; k < M; ++k) {
        // row k at the wave's points, lanes = expressions (all loads issued
        // before the first use)
        T gv[PW], vv[PW];
#pragma unroll
        for (int i = 0; i < PW; ++i) {
            const int64_t n = n0 + wave + i * GA_WAVES;
            const bool ok = n < N && e < E;
            const int64_t off = ok ? ((int64_t)k * N + n) * E + e : 0;
            gv[i] = ok ? grad_cache[off] : T(0);
            if constexpr (VALS) vv[i] = (ok && k < myL) ? cache[off] : T(0);
        }
#pragma unroll
        for (int i = 0; i < PW; ++i) {
            const int pp = wave + i * GA_WAVES;
            if (k < myL) {
                gtile[lane][pp] = gv[i];
                if constexpr (VALS) vtile[lane][pp] = vv[i];
            } else if (gv[i] != T(0)) {
                lerr = ERR_GRAD_NO_OP;
            }
        }
        if (k < Lmax) {
            __syncthreads();
            const int64_t n = n0 + lane;
            for (int q = wave; q < FG; q += GA_WAVES) {
                if (k < L_s[q] && n < N) {
                    const int64_t o = base_s[q] + (int64_t)k * N + n;
                    grows[o] = gtile[q][lane];
                    if constexpr (VALS) vrows[o] = vtile[q][lane];
                }
            }
            __syncthreads();
        }
    }
    lerr = wave_max_i32(lerr);
    if (lane == 0 && lerr) atomicMax(err_flag, lerr);
}

// ROWS: the blocks whose group was gathered (fr), with the gathered value
// and gradient rows as the tapes (the adjoints accumulate in place in the
// gradient rows, a copy), so the tapes take no LDS and the kernel runs at
// full occupancy (small LDS tapes for short programs, 32 KB a block, were
// slower); otherwise the other blocks, with LDS tapes loaded from the cache
// layout.
template <typename T, bool ROWS, bool MANYX>
__global__ __launch_bounds__(BWD_WAVES * 64) __attribute__((amdgpu_waves_per_eu(ROWS && !MANYX && sizeof(T) == 4 ? BWR_WPE : 1))) void
k_full_backward_grp(
    const T *__restrict__ grad_cache, const T *__restrict__ cache, const int64_t *__restrict__ ops,
    const int64_t *__restrict__ ch, int M, int64_t B, int64_t K, int64_t N, int n_x, T *__restrict__ grad_C,
    T *__restrict__ gx_partial, int32_t *err_flag, FwdRows fr,
    const T *__restrict__ vrows, int64_t blk_begin) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ int Ls[BWD_WAVES];
    const int wave = uniform_i32((int)(threadIdx.x / WAVE));
    const int lane = lane_id();
    const size_t tstride = (size_t)M * WAVE + 1;
    const size_t tb = ROWS ? 0 : (size_t)BWD_WAVES * tstride;  // elements of one block's tapes
    T *vals = reinterpret_cast<T *>(smem);
    T *adjs = vals + tb;
    double *cextra = reinterpret_cast<double *>(adjs + tb + (tb * sizeof(T) % 8 ? 1 : 0)) + (size_t)wave * M;
    T *gxs = reinterpret_cast<T *>(reinterpret_cast<double *>(adjs + tb + (tb * sizeof(T) % 8 ? 1 : 0)) +
                                   (size_t)BWD_WAVES * M);  // [4][64][NXREG]
    const int64_t E = B * K;
    // XCD-remapped so that neighbouring expression groups share one XCD's L2
    const int64_t blk = blk_begin + xcd_remap(blockIdx.x, (int64_t)gridDim.x);
    const int64_t e0 = blk * BWD_WAVES;
    // the block's expressions lie in one row-buffer group (BWD_WAVES divides FG)
    const int64_t grp = e0 / FG;
    const bool from_rows = fr.cap >= 0 && fr.goff[grp + 1] <= fr.cap;  // uniform over the block
    if (from_rows != ROWS) return;  // the other launch handles this block
    const int64_t e_id = e0 + wave;
    const bool bvalid = e_id < E;
    const int64_t b = bvalid ? e_id / K : 0, j = bvalid ? e_id % K : 0;
    const bool want_gx = gx_partial != nullptr;
    Program<T> p;
    if (bvalid) p = load_program<T>(ops, ch, (const T *)nullptr, b, 1, B, 1, B, M, n_x);
    else { p.L = 0; p.w0 = p.w1 = 0; p.cmask0 = p.cmask1 = 0; p.root = 0; p.verr = 0; p.c0 = p.c1 = T(0); }
    for (int k = lane; k < M; k += WAVE) cextra[k] = 0.0;
    if (lane == 0) Ls[wave] = p.L;
    GradAcc<T> acc;
#pragma unroll
    for (int q = 0; q < NCREG; ++q) acc.cacc[q] = 0.0;
    int32_t err = 0, lerr = 0;
    T *vt = vals + (size_t)wave * tstride;
    T *at = adjs + (size_t)wave * tstride;
    const int t = (int)threadIdx.x;
    const int el = t % BWD_WAVES, nl = t / BWD_WAVES;  // nl < 64
    const int64_t el_e = e0 + el;
    const int64_t n_tiles = (N + TILE - 1) / TILE;
    const int64_t rbase = (from_rows && bvalid) ? fr.goff[grp] + fr.eoff[e_id] : 0;
    const int Lr = (from_rows && bvalid) ? (p.L < (int)fr.Lw[e_id] ? p.L : (int)fr.Lw[e_id]) : 0;
    __syncthreads();
    const int Lel = Ls[el];
    if constexpr (ROWS && !MANYX) {
        // BWR_U tiles at a time, node by node (backward_rows_batch)
        T *gr0 = reinterpret_cast<T *>(fr.rows) + rbase;
        const T *vr0 = vrows + rbase;
        for (int64_t t0 = 0; t0 < n_tiles; t0 += BWR_U) {
            int64_t nu[BWR_U];
            bool au[BWR_U];
#pragma unroll
            for (int u = 0; u < BWR_U; ++u) {
                const int64_t n = (t0 + u) * TILE + lane;
                au[u] = bvalid && t0 + u < n_tiles && n < N;
                nu[u] = au[u] ? n : 0;
            }
            if (Lr > 0) {
                const int32_t e = backward_rows_batch<T, BWR_U>(p, vr0, gr0, N, Lr, nu, au, acc, cextra);
                err = e > err ? e : err;
            }
            if (want_gx) {
#pragma unroll
                for (int u = 0; u < BWR_U; ++u) {
                    if (t0 + u >= n_tiles) break;  // uniform over the block
                    T gxl[NXREG];
                    auto adj = [&](int k) -> T & { return gr0[(int64_t)(k < Lr ? k : Lr - 1) * N + nu[u]]; };
                    var_adjoints<T>(p, adj, Lr, au[u], 0, gxl);
#pragma unroll
                    for (int q = 0; q < NXREG; ++q) gxs[(wave * WAVE + lane) * NXREG + q] = gxl[q];
                    __syncthreads();
                    const int64_t n = (t0 + u) * TILE + lane;
                    if (wave == 0 && n < N) {
                        for (int v = 0; v < n_x; ++v) {
                            T sum = gxs[(0 * WAVE + lane) * NXREG + v];
                            for (int w2 = 1; w2 < BWD_WAVES; ++w2) sum += gxs[(w2 * WAVE + lane) * NXREG + v];
                            gx_partial[blk * N * n_x + n * n_x + v] = sum;
                        }
                    }
                    __syncthreads();
                }
            }
        }
    } else
    for (int64_t tt = 0; tt < n_tiles; ++tt) {
        if constexpr (!ROWS) {
        // cooperative load, expression index fastest
        {
            const int64_t n = tt * TILE + nl;
            const bool ok = el_e < E && n < N;
            T *vd = vals + (size_t)el * tstride + nl;
            T *ad = adjs + (size_t)el * tstride + nl;
            // BWD_BATCH rows of loads in flight before they are used
            for (int k0 = 0; k0 < M; k0 += BWD_BATCH) {
                T g[BWD_BATCH], v[BWD_BATCH];
#pragma unroll
                for (int u = 0; u < BWD_BATCH; ++u) {
                    const int k = k0 + u;
                    const int64_t off = ((int64_t)k * N + n) * E + el_e;
                    g[u] = (ok && k < M) ? grad_cache[off] : T(0);
                    v[u] = (ok && k < Lel) ? cache[off] : T(0);
                }
#pragma unroll
                for (int u = 0; u < BWD_BATCH; ++u) {
                    const int k = k0 + u;
                    if (k < Lel) {
                        ad[(size_t)k * WAVE] = g[u];
                        vd[(size_t)k * WAVE] = v[u];
                    } else if (g[u] != T(0)) {
                        lerr = ERR_GRAD_NO_OP;
                    }
                }
            }
        }
        __syncthreads();  // the LDS tapes are loaded (the row tapes need none)
        }
        const int64_t n = tt * TILE + lane;
        const bool active = bvalid && n < N;
        const int64_t nn = active ? n : 0;
#pragma unroll
        for (int q = 0; q < NXREG; ++q) acc.gxl[q] = T(0);
        int32_t e;
        if constexpr (ROWS) {
            // rows clamped below Lr (a valid program only names children
            // below their parent; an unvalidated one stays in bounds)
            T *gr = reinterpret_cast<T *>(fr.rows) + rbase + nn;
            const T *vr = vrows + rbase + nn;
            auto val = [&](int k) -> T { return vr[(int64_t)(k < Lr ? k : Lr - 1) * N]; };
            auto adj = [&](int k) -> T & { return gr[(int64_t)(k < Lr ? k : Lr - 1) * N]; };
            e = backward_tile<T, true>(p, val, adj, Lr, active, acc, cextra, n_x);
            if constexpr (MANYX)
                if (want_gx)
                    flush_gx<T, MANYX>(p, adj, Lr, active, n_x, acc.gxl, gxs, wave, lane, n, N,
                                       gx_partial + blk * N * n_x);
        } else {
            auto val = [&](int k) -> T { return vt[(size_t)k * WAVE + lane]; };
            auto adj = [&](int k) -> T & { return at[(size_t)k * WAVE + lane]; };
            e = backward_tile<T, true>(p, val, adj, p.L, active, acc, cextra, n_x);
            if constexpr (MANYX)
                if (want_gx)
                    flush_gx<T, MANYX>(p, adj, p.L, active, n_x, acc.gxl, gxs, wave, lane, n, N,
                                       gx_partial + blk * N * n_x);
        }
        err = e > err ? e : err;
        if constexpr (!MANYX) {
            if (want_gx) {
#pragma unroll
                for (int q = 0; q < NXREG; ++q) gxs[(wave * WAVE + lane) * NXREG + q] = acc.gxl[q];
                __syncthreads();
                if (wave == 0 && n < N) {
                    for (int v = 0; v < n_x; ++v) {
                        T sum = gxs[(0 * WAVE + lane) * NXREG + v];
                        for (int w2 = 1; w2 < BWD_WAVES; ++w2) sum += gxs[(w2 * WAVE + lane) * NXREG + v];
                        gx_partial[blk * N * n_x + n * n_x + v] = sum;
                    }
                }
            }
        }
        __syncthreads();  // LDS tapes are reloaded for the next tile
    }
    if (lerr) atomicMax(err_flag, lerr);
    if (!bvalid) return;
    double lo, hi;
    finish_constants<T>(p, acc, cextra, lo, hi);
    if (lane < M) grad_C[((int64_t)lane * B + b) * K + j] = (T)lo;
    if (lane + WAVE < M) grad_C[((int64_t)(lane + WAVE) * B + b) * K + j] = (T)hi;
    err = wave_max_i32(err);
    if (lane == 0 && err) atomicMax(err_flag, err);
}

template <typename T>
inline size_t backward_grp_lds(int64_t M, bool rows = false) {
    const size_t tapes = rows ? 0 : 2 * (size_t)BWD_WAVES * ((size_t)M * WAVE + 1) * sizeof(T);
    return (tapes + 8) + (size_t)BWD_WAVES * M * sizeof(double) + (size_t)BWD_WAVES * WAVE * NXREG * sizeof(T);
}

// Sums P partial rows of length len in FP64, in row order.
template <typename T>
__global__ void k_reduce_partials(const T *__restrict__ part, int64_t P, int64_t len, T *__restrict__ out,
                                  int accumulate) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= len) return;
    out[i] = (T)sum_rows_in_order(part + i, 0, P, len, accumulate ? (double)out[i] : 0.0);
}

// The same sum in two stages (the single-stage kernel gives each column one
// thread, so a few thousand columns over thousands of rows leave the GPU
// almost idle): stage 1 sums chunks of RED_ROWS rows per column in FP64,
// stage 2 adds the chunk sums in chunk order.  Deterministic.
constexpr int64_t RED_ROWS = 64;
template <typename T>
__global__ void k_reduce_partials_1(const T *__restrict__ part, int64_t P, int64_t len, double *__restrict__ part64) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= len) return;
    const int64_t r0 = (int64_t)blockIdx.y * RED_ROWS, r1 = r0 + RED_ROWS < P ? r0 + RED_ROWS : P;
    part64[(int64_t)blockIdx.y * len + i] = sum_rows_in_order(part + i, r0, r1, len, 0.0);
}
template <typename T>
__global__ void k_reduce_partials_2(const double *__restrict__ part64, int64_t chunks, int64_t len,
                                    T *__restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= len) return;
    out[i] = (T)sum_rows_in_order(part64 + i, 0, chunks, len, 0.0);
}

inline int ret_last() { return (int)hipGetLastError(); }

template <typename T>
int forward_mb(const T *X, int64_t N, int64_t n_x, const int64_t *ops, const int64_t *ch, const T *C, int64_t M,
               int64_t B, int64_t K, T *cache, void *ws, int64_t ws_bytes, hipStream_t s) {
    if (M <= 0 || M > 128 || N < 0 || B < 0 || n_x < 0 || K <= 0 || ws_bytes < 0) return -1;
    if (N == 0 || B == 0) return 0;
    int64_t n_tiles = (N + TILE - 1) / TILE;
    int64_t nblk = n_tiles * B * K;
    if (nblk > 0x7fffffffLL) return -1;
    // row buffer (two-pass forward) when the workspace holds the index arrays
    const int64_t E = B * K, G = (E + FG - 1) / FG, ib = fwd_index_bytes(E);
    FwdRows fr{nullptr, nullptr, nullptr, nullptr, -1};
    if (ws != nullptr && ws_bytes >= ib + 256 && G * n_tiles <= 0x7fffffffLL) {
        unsigned char *w = static_cast<unsigned char *>(ws);
        const int64_t o1 = align256(8 * (G + 1)), o2 = o1 + align256(8 * E);
        fr.goff = reinterpret_cast<const int64_t *>(w);
        fr.eoff = reinterpret_cast<const int64_t *>(w + o1);
        fr.Lw = w + o2;
        fr.rows = w + ib;
        fr.cap = (ws_bytes - ib) / (int64_t)sizeof(T);
    }
    // grouped kernel: the most expressions per block whose tapes fit in LDS
    for (int EW = 8; EW >= 2; EW /= 2) {
        const size_t lds = (size_t)EW * ((size_t)M * WAVE + 1) * sizeof(T);
        if (lds > lds_limit_device()) continue;
        const int64_t groups = (B * K + EW - 1) / EW;
        const int64_t nb = n_tiles * groups;
        if (nb > 0x7fffffffLL) break;
        if (fr.cap >= 0) {
            hipLaunchKernelGGL(k_fwd_plan, dim3((unsigned)G), dim3(64), 0, s, ops, (int)M, B, K, N,
                               const_cast<int64_t *>(fr.goff), const_cast<int64_t *>(fr.eoff),
                               const_cast<uint8_t *>(fr.Lw));
            hipLaunchKernelGGL(k_fwd_scan, dim3(1), dim3(1024), 0, s, const_cast<int64_t *>(fr.goff), G);
            // (run one after the other: concurrent chunks of the two passes
            // slow each other down more than they overlap)
            const int64_t nrb = (E * ((n_tiles + FR_TILES - 1) / FR_TILES) + FR_WAVES - 1) / FR_WAVES;
            hipLaunchKernelGGL(k_fwd_rows<T>, dim3((unsigned)nrb), dim3(FR_WAVES * 64), 0, s, X, N, (int)n_x, ops, ch, C,
                               (int)M, B, K, n_tiles, (int64_t)0, E, fr, cache);
            hipLaunchKernelGGL(k_fwd_transpose<T>, dim3((unsigned)(G * n_tiles)), dim3(TR_WAVES * 64), 0, s, (int)M, N,
                               E, (int64_t)0, G, fr, cache);
            return ret_last();
        }
        hipLaunchKernelGGL(k_full_forward_grp<T>, dim3((unsigned)nb), dim3(64 * EW), lds, s, X, N, (int)n_x, ops, ch,
                           C, (int)M, B, K, n_tiles, EW, cache);
        return ret_last();
    }
    size_t lds = (size_t)M * WAVE * sizeof(T);
    if (lds > lds_limit()) return -3;
    hipLaunchKernelGGL(k_full_forward<T>, dim3((unsigned)nblk), dim3(64), lds, s, X, N, (int)n_x, ops, ch, C,
                       (int)M, B, K, n_tiles, cache);
    return ret_last();
}

template <typename T>
int64_t backward_mb_ws_bytes(int64_t B, int64_t N, int64_t n_x, int64_t K = 1) {
    // per-block partial rows of grad_X, then the FP64 chunk sums of the reduction
    const int64_t nblk = (B * K + BWD_WAVES - 1) / BWD_WAVES;
    const int64_t part = (nblk * N * n_x * (int64_t)sizeof(T) + 255) / 256 * 256;
    const int64_t chunks = (nblk + RED_ROWS - 1) / RED_ROWS;
    return part + chunks * N * n_x * 8 + 256;
}

// X and C (both or neither): the gathered groups recompute their forward
// values from them instead of reading the cache (which must then be their
// forward output); the other groups still read the cache.
template <typename T>
int backward_mb(T *grad_out, const T *cache, const int64_t *ops, const int64_t *ch, int64_t M, int64_t B,
                int64_t K, int64_t N, int64_t n_x, T *grad_C, T *grad_X, void *workspace, int64_t ws_bytes,
                int32_t *d_err, hipStream_t s, void *rows_ws = nullptr, int64_t rows_bytes = 0,
                const T *X = nullptr, const T *C = nullptr) {
    if (M <= 0 || M > 128 || N < 0 || B < 0 || n_x < 0 || K <= 0 || rows_bytes < 0) return -1;
    if ((X == nullptr) != (C == nullptr)) return -1;
    const bool recomp = X != nullptr;
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, d_err, 0, (int64_t)1);
    int64_t nblk = (B * K + BWD_WAVES - 1) / BWD_WAVES;
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
    if (grad_X) {
        if (ws_bytes < backward_mb_ws_bytes<T>(B, N, n_x, K)) return -2;
        gx_partial = (T *)workspace;
    }
    if (backward_grp_lds<T>(M) <= lds_limit_device()) {
        // gradient and value rows of each group through the row workspace
        // (half each), when it holds the index arrays
        const int64_t E = B * K, G = (E + FG - 1) / FG, ib = fwd_index_bytes(E);
        const int64_t n_tiles = (N + TILE - 1) / TILE;
        FwdRows fr{nullptr, nullptr, nullptr, nullptr, -1};
        T *vrows = nullptr;
        const bool manyx = n_x > NXREG;
        if (rows_ws != nullptr && rows_bytes >= ib + 512 && G * n_tiles <= 0x7fffffffLL) {
            unsigned char *w = static_cast<unsigned char *>(rows_ws);
            const int64_t o1 = align256(8 * (G + 1)), o2 = o1 + align256(8 * E);
            const int64_t half = (rows_bytes - ib) / 2 / 256 * 256;
            fr.goff = reinterpret_cast<const int64_t *>(w);
            fr.eoff = reinterpret_cast<const int64_t *>(w + o1);
            fr.Lw = w + o2;
            fr.rows = w + ib;
            fr.cap = half / (int64_t)sizeof(T);
            vrows = reinterpret_cast<T *>(w + ib + half);
            hipLaunchKernelGGL(k_fwd_plan, dim3((unsigned)G), dim3(64), 0, s, ops, (int)M, B, K, N,
                               const_cast<int64_t *>(fr.goff), const_cast<int64_t *>(fr.eoff),
                               const_cast<uint8_t *>(fr.Lw));
            hipLaunchKernelGGL(k_fwd_scan, dim3(1), dim3(1024), 0, s, const_cast<int64_t *>(fr.goff), G);
            // The gather (and, recomputing, the forward's row pass that
            // writes the value rows) on two side streams at once, then the
            // reverse sweep on the caller's stream, chunk by chunk of groups.
            // One chunk is fastest at 65536 x 1024 (10.8 ms): with 8 chunks,
            // the sweep of one chunk under the gather of the next, the kernels
            // slowed each other down (14.3 ms).
            const bool rec_ok = recomp && (E * ((n_tiles + FR_TILES - 1) / FR_TILES) + FR_WAVES - 1) / FR_WAVES <=
                                              0x7fffffffLL;
            const int64_t nch = G < BWD_CHUNKS ? G : BWD_CHUNKS;
            const int64_t gpc = (G + nch - 1) / nch;      // groups per chunk
            constexpr int64_t BPG = FG / BWD_WAVES;       // sweep blocks per group
            ForkScope fk(s);
            int r = fk.fork();
            if (r) return fk.finish(r);
            for (int64_t g0 = 0; g0 < G; g0 += gpc) {
                const int64_t gc = g0 + gpc < G ? gpc : G - g0;
                if (rec_ok) {
                    hipLaunchKernelGGL((k_bwd_gather<T, false>), dim3((unsigned)(gc * n_tiles)), dim3(GA_WAVES * 64),
                                       0, fk.side(0), grad_out, cache, (int)M, N, E, g0, gc, fr, vrows, d_err);
                    FwdRows frv = fr;
                    frv.rows = vrows;
                    const int64_t e0 = g0 * FG, ec = (g0 + gc) * FG < E ? gc * FG : E - e0;
                    const int64_t nrb = (ec * ((n_tiles + FR_TILES - 1) / FR_TILES) + FR_WAVES - 1) / FR_WAVES;
                    hipLaunchKernelGGL(k_fwd_rows<T>, dim3((unsigned)nrb), dim3(FR_WAVES * 64), 0, fk.side(1), X, N,
                                       (int)n_x, ops, ch, C, (int)M, B, K, n_tiles, e0, ec, frv, (T *)nullptr);
                } else {
                    hipLaunchKernelGGL((k_bwd_gather<T, true>), dim3((unsigned)(gc * n_tiles)), dim3(GA_WAVES * 64),
                                       0, fk.side(0), grad_out, cache, (int)M, N, E, g0, gc, fr, vrows, d_err);
                }
                if ((r = (int)hipGetLastError())) return fk.finish(r);
                if ((r = fk.handoff(0))) return fk.finish(r);
                if (rec_ok && (r = fk.handoff(1))) return fk.finish(r);
                const int64_t b0 = g0 * BPG, b1 = (g0 + gc) * BPG < nblk ? (g0 + gc) * BPG : nblk;
                hipLaunchKernelGGL(manyx ? (k_full_backward_grp<T, true, true>) : (k_full_backward_grp<T, true, false>),
                                   dim3((unsigned)(b1 - b0)), dim3(BWD_WAVES * 64), backward_grp_lds<T>(M, true), s,
                                   grad_out, cache, ops, ch, (int)M, B, K, N, (int)n_x, grad_C, gx_partial, d_err, fr,
                                   (const T *)vrows, b0);
                if ((r = (int)hipGetLastError())) return fk.finish(r);
            }
            if ((r = fk.finish(0))) return r;
        }
        // the others (all of them without a row workspace)
        hipLaunchKernelGGL(manyx ? (k_full_backward_grp<T, false, true>) : (k_full_backward_grp<T, false, false>),
                           dim3((unsigned)nblk), dim3(BWD_WAVES * 64), backward_grp_lds<T>(M), s, grad_out, cache,
                           ops, ch, (int)M, B, K, N, (int)n_x, grad_C, gx_partial, d_err, fr, (const T *)vrows,
                           (int64_t)0);
    } else {
        size_t lds = (size_t)BWD_WAVES * M * sizeof(double) + (size_t)BWD_WAVES * WAVE * NXREG * sizeof(T);
        if (lds > lds_limit()) return -3;
        hipLaunchKernelGGL(n_x > NXREG ? (k_full_backward<T, true>) : (k_full_backward<T, false>), dim3((unsigned)nblk),
                           dim3(BWD_WAVES * 64), lds, s, grad_out, cache, ops, ch, (int)M, B, K, N, (int)n_x, grad_C,
                           gx_partial, d_err);
    }
    if (gx_partial) {
        const int64_t len = N * n_x;
        const int64_t chunks = (nblk + RED_ROWS - 1) / RED_ROWS;
        const int64_t part = (nblk * len * (int64_t)sizeof(T) + 255) / 256 * 256;
        double *part64 = reinterpret_cast<double *>(static_cast<unsigned char *>(workspace) + part);
        if (chunks <= 65535) {
            hipLaunchKernelGGL(k_reduce_partials_1<T>, dim3((unsigned)((len + 255) / 256), (unsigned)chunks), dim3(256),
                               0, s, (const T *)gx_partial, nblk, len, part64);
            hipLaunchKernelGGL(k_reduce_partials_2<T>, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, s,
                               (const double *)part64, chunks, len, grad_X);
        } else {
            hipLaunchKernelGGL(k_reduce_partials<T>, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, s,
                               (const T *)gx_partial, nblk, len, grad_X, 0);
        }
    }
    return ret_last();
}

}  // namespace

extern "C" {

// 1 when sr_backward_mb_* of this size only reads grad_cache (the grouped
// kernel), 0 when it uses grad_cache as scratch.
int sr_backward_mb_reads_only(int64_t M, int is_f64) {
    if (M <= 0 || M > 128) return 0;
    return (is_f64 ? backward_grp_lds<double>(M) : backward_grp_lds<float>(M)) <= lds_limit_device() ? 1 : 0;
}

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
    return forward_mb<float>(X, N, n_x, ops, ch, C, M, B, 1, cache, nullptr, 0, (hipStream_t)stream);
}
int sr_forward_mb_f64(const double *X, int64_t N, int64_t n_x, const int64_t *ops, const int64_t *ch,
                      const double *C, int64_t M, int64_t B, double *cache, void *stream) {
    return forward_mb<double>(X, N, n_x, ops, ch, C, M, B, 1, cache, nullptr, 0, (hipStream_t)stream);
}

int64_t sr_forward_mb_workspace_bytes(int64_t M, int64_t N, int64_t B, int64_t K, int is_f64) {
    if (M <= 0 || N <= 0 || B <= 0 || K <= 0) return 0;
    const int64_t E = B * K;
    const int64_t rows = (M * N * E + FWD_ROWS_SHARE - 1) / FWD_ROWS_SHARE;  // a quarter of the cache
    return fwd_index_bytes(E) + align256(rows * (is_f64 ? 8 : 4));
}
int sr_forward_mb_ws_f32(const float *X, int64_t N, int64_t n_x, const int64_t *ops, const int64_t *ch,
                         const float *C, int64_t M, int64_t B, float *cache, void *workspace, int64_t ws_bytes,
                         void *stream) {
    return forward_mb<float>(X, N, n_x, ops, ch, C, M, B, 1, cache, workspace, ws_bytes, (hipStream_t)stream);
}
int sr_forward_mb_ws_f64(const double *X, int64_t N, int64_t n_x, const int64_t *ops, const int64_t *ch,
                         const double *C, int64_t M, int64_t B, double *cache, void *workspace, int64_t ws_bytes,
                         void *stream) {
    return forward_mb<double>(X, N, n_x, ops, ch, C, M, B, 1, cache, workspace, ws_bytes, (hipStream_t)stream);
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

int64_t sr_backward_mb_rows_workspace_bytes(int64_t M, int64_t N, int64_t B, int64_t K, int is_f64) {
    if (M <= 0 || N <= 0 || B <= 0 || K <= 0) return 0;
    const int64_t E = B * K;
    const int64_t rows = (M * N * E + FWD_ROWS_SHARE - 1) / FWD_ROWS_SHARE;  // a quarter of the cache, twice
    return fwd_index_bytes(E) + 2 * align256(rows * (is_f64 ? 8 : 4));
}
int sr_backward_mb_rows_f32(float *grad_out, const float *cache, const int64_t *ops, const int64_t *ch, int64_t M,
                            int64_t B, int64_t N, int64_t n_x, float *grad_C, float *grad_X, void *workspace,
                            int64_t ws_bytes, void *rows_ws, int64_t rows_bytes, int32_t *d_err, void *stream) {
    return backward_mb<float>(grad_out, cache, ops, ch, M, B, 1, N, n_x, grad_C, grad_X, workspace, ws_bytes, d_err,
                              (hipStream_t)stream, rows_ws, rows_bytes);
}
int sr_backward_mb_rows_f64(double *grad_out, const double *cache, const int64_t *ops, const int64_t *ch, int64_t M,
                            int64_t B, int64_t N, int64_t n_x, double *grad_C, double *grad_X, void *workspace,
                            int64_t ws_bytes, void *rows_ws, int64_t rows_bytes, int32_t *d_err, void *stream) {
    return backward_mb<double>(grad_out, cache, ops, ch, M, B, 1, N, n_x, grad_C, grad_X, workspace, ws_bytes, d_err,
                               (hipStream_t)stream, rows_ws, rows_bytes);
}

// The gathered groups recompute their forward values from X and C (the
// inputs of the forward that wrote `cache`) instead of reading the cache.
int sr_backward_mb_rows_recompute_f32(float *grad_out, const float *cache, const float *X, const float *C,
                                      const int64_t *ops, const int64_t *ch, int64_t M, int64_t B, int64_t N,
                                      int64_t n_x, float *grad_C, float *grad_X, void *workspace, int64_t ws_bytes,
                                      void *rows_ws, int64_t rows_bytes, int32_t *d_err, void *stream) {
    if (X == nullptr || C == nullptr) return -1;
    return backward_mb<float>(grad_out, cache, ops, ch, M, B, 1, N, n_x, grad_C, grad_X, workspace, ws_bytes, d_err,
                              (hipStream_t)stream, rows_ws, rows_bytes, X, C);
}
int sr_backward_mb_rows_recompute_f64(double *grad_out, const double *cache, const double *X, const double *C,
                                      const int64_t *ops, const int64_t *ch, int64_t M, int64_t B, int64_t N,
                                      int64_t n_x, double *grad_C, double *grad_X, void *workspace, int64_t ws_bytes,
                                      void *rows_ws, int64_t rows_bytes, int32_t *d_err, void *stream) {
    if (X == nullptr || C == nullptr) return -1;
    return backward_mb<double>(grad_out, cache, ops, ch, M, B, 1, N, n_x, grad_C, grad_X, workspace, ws_bytes, d_err,
                               (hipStream_t)stream, rows_ws, rows_bytes, X, C);
}

int sr_multi_forward_mb_f32(const float *X, int64_t N, int64_t n_x, const int64_t *ops, const int64_t *ch,
                            const float *C, int64_t M, int64_t B, int64_t K, float *cache, void *stream) {
    return forward_mb<float>(X, N, n_x, ops, ch, C, M, B, K, cache, nullptr, 0, (hipStream_t)stream);
}
int sr_multi_forward_mb_f64(const double *X, int64_t N, int64_t n_x, const int64_t *ops, const int64_t *ch,
                            const double *C, int64_t M, int64_t B, int64_t K, double *cache, void *stream) {
    return forward_mb<double>(X, N, n_x, ops, ch, C, M, B, K, cache, nullptr, 0, (hipStream_t)stream);
}
int sr_multi_forward_mb_ws_f32(const float *X, int64_t N, int64_t n_x, const int64_t *ops, const int64_t *ch,
                               const float *C, int64_t M, int64_t B, int64_t K, float *cache, void *workspace,
                               int64_t ws_bytes, void *stream) {
    return forward_mb<float>(X, N, n_x, ops, ch, C, M, B, K, cache, workspace, ws_bytes, (hipStream_t)stream);
}
int sr_multi_forward_mb_ws_f64(const double *X, int64_t N, int64_t n_x, const int64_t *ops, const int64_t *ch,
                               const double *C, int64_t M, int64_t B, int64_t K, double *cache, void *workspace,
                               int64_t ws_bytes, void *stream) {
    return forward_mb<double>(X, N, n_x, ops, ch, C, M, B, K, cache, workspace, ws_bytes, (hipStream_t)stream);
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
int sr_multi_backward_mb_rows_f32(float *grad_out, const float *cache, const int64_t *ops, const int64_t *ch,
                                  int64_t M, int64_t B, int64_t N, int64_t n_x, int64_t K, float *grad_C,
                                  float *grad_X, void *workspace, int64_t ws_bytes, void *rows_ws, int64_t rows_bytes,
                                  int32_t *d_err, void *stream) {
    return backward_mb<float>(grad_out, cache, ops, ch, M, B, K, N, n_x, grad_C, grad_X, workspace, ws_bytes, d_err,
                              (hipStream_t)stream, rows_ws, rows_bytes);
}
int sr_multi_backward_mb_rows_f64(double *grad_out, const double *cache, const int64_t *ops, const int64_t *ch,
                                  int64_t M, int64_t B, int64_t N, int64_t n_x, int64_t K, double *grad_C,
                                  double *grad_X, void *workspace, int64_t ws_bytes, void *rows_ws,
                                  int64_t rows_bytes, int32_t *d_err, void *stream) {
    return backward_mb<double>(grad_out, cache, ops, ch, M, B, K, N, n_x, grad_C, grad_X, workspace, ws_bytes, d_err,
                               (hipStream_t)stream, rows_ws, rows_bytes);
}
int sr_multi_backward_mb_rows_recompute_f32(float *grad_out, const float *cache, const float *X, const float *C,
                                            const int64_t *ops, const int64_t *ch, int64_t M, int64_t B, int64_t N,
                                            int64_t n_x, int64_t K, float *grad_C, float *grad_X, void *workspace,
                                            int64_t ws_bytes, void *rows_ws, int64_t rows_bytes, int32_t *d_err,
                                            void *stream) {
    if (X == nullptr || C == nullptr) return -1;
    return backward_mb<float>(grad_out, cache, ops, ch, M, B, K, N, n_x, grad_C, grad_X, workspace, ws_bytes, d_err,
                              (hipStream_t)stream, rows_ws, rows_bytes, X, C);
}
int sr_multi_backward_mb_rows_recompute_f64(double *grad_out, const double *cache, const double *X, const double *C,
                                            const int64_t *ops, const int64_t *ch, int64_t M, int64_t B, int64_t N,
                                            int64_t n_x, int64_t K, double *grad_C, double *grad_X, void *workspace,
                                            int64_t ws_bytes, void *rows_ws, int64_t rows_bytes, int32_t *d_err,
                                            void *stream) {
    if (X == nullptr || C == nullptr) return -1;
    return backward_mb<double>(grad_out, cache, ops, ch, M, B, K, N, n_x, grad_C, grad_X, workspace, ws_bytes, d_err,
                               (hipStream_t)stream, rows_ws, rows_bytes, X, C);
}

}  // extern "C"
