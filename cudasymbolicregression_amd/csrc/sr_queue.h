// Deterministic bucket lists for the persistent work queues.
//
// Items (expressions or fitting problems) are classified into NB buckets
// (the caller's kernel writes each item's bucket, one wave per item, since
// a program row is read by a whole wave), then in runs of 64 consecutive
// items, one wave per run:
//   1. the wave counts its run's items per bucket;
//   2. one block turns the counts into per-run offsets (exclusive prefix
//      sums over runs, bucket by bucket) and the bucket totals;
//   3. each wave scatters its run into the bucket lists at those offsets.
// Every bucket list is in item order, the same on every call, with no
// atomics.  Workspace: bucket_of [B] int8, run counts and offsets [R][NB]
// int32 each, R = ceil(B / 64).
//
// Static segments (optional).  When the caller also gives each item a cost
// (cost_of [B] int32, > 0) and the number of waves W_k that will work on
// bucket k, the same passes cut every bucket list into W_k contiguous
// segments of nearly equal cost: wave g takes the items whose exclusive
// cost prefix c0 lies in [g T / W, (g+1) T / W), T the bucket's total.  The
// segments are written as seg[k][g] = first list position of wave g
// (g = 0..W_k; seg[k][W_k] = the bucket's count).  Item j owns the
// boundaries g T / W in (c0_j, c0_j + cost_j] and sets their starts to
// j + 1; boundary 0 is position 0.  A kernel that takes its work from the
// segments instead of a shared counter processes the same items in the same
// order on every call, which is what deterministic per-wave partial sums
// need.  Extra workspace: run costs and their offsets [R][NB] int64 each,
// bucket cost totals [NB] int64.
#pragma once

#include "sr_wave.h"
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sr {

constexpr int QRUN = 64;    // items per classification run (one wave)
constexpr int QNB_MAX = 8;  // most buckets a queue serves

inline int64_t queue_runs(int64_t B) { return (B + QRUN - 1) / QRUN; }
inline size_t queue_ws_bytes(int64_t B, int NB) {
    return (((size_t)B + 255) / 256 * 256) + 2 * (size_t)queue_runs(B) * NB * sizeof(int32_t) + 256;
}
// Extra workspace of the static segments (cost_of, run costs, offsets, totals).
inline size_t queue_seg_ws_bytes(int64_t B, int NB) {
    return ((size_t)B * sizeof(int32_t) + 255) / 256 * 256 + 2 * (size_t)queue_runs(B) * NB * sizeof(int64_t) +
           (size_t)NB * sizeof(int64_t) + 256;
}

// Where the segments go: seg[k] has W[k] + 1 entries.
struct QueueSegs {
    int32_t *seg[QNB_MAX];
    int64_t W[QNB_MAX];
};

namespace {

// Pass 2: one block of 256 threads; thread t owns a contiguous range of runs.
// The 256 thread totals are scanned per wave with shuffles, then across the
// 4 wave totals (integer sums: the offsets do not depend on the order).
template <typename V>
__device__ inline void scan_runs(const V *__restrict__ runval, int64_t R, int NB, V *__restrict__ runoff,
                                 V *__restrict__ total, V *wtot) {
    const int t = threadIdx.x, lane = lane_id(), w = t / WAVE;
    const int64_t per = (R + 255) / 256;
    const int64_t r0 = t * per, r1 = r0 + per < R ? r0 + per : R;
    for (int k = 0; k < NB; ++k) {
        V s = 0;
        for (int64_t r = r0; r < r1; ++r) s += runval[r * NB + k];
        V incl = s;
#pragma unroll
        for (int d = 1; d < WAVE; d <<= 1) {
            const V up = __shfl_up(incl, d, WAVE);
            if (lane >= d) incl += up;
        }
        if (lane == WAVE - 1) wtot[w] = incl;
        __syncthreads();
        V before = 0;
        for (int q = 0; q < w; ++q) before += wtot[q];
        if (t == 255) total[k] = before + incl;
        V acc = before + incl - s;
        for (int64_t r = r0; r < r1; ++r) {
            runoff[r * NB + k] = acc;
            acc += runval[r * NB + k];
        }
        __syncthreads();  // wtot is rewritten for the next bucket
    }
}

// With segments it also writes each segment list's entries that pass 3 does
// not: seg[k][0] = 0, and every entry of an empty bucket's list (pass 3
// writes entries 1..W[k] of a non-empty one, whose items cover every
// boundary).
__global__ __launch_bounds__(256) void k_queue_scan(const int32_t *__restrict__ runcnt, int64_t R, int NB,
                                                   int32_t *__restrict__ runoff, int32_t *__restrict__ counts,
                                                   const int64_t *__restrict__ runcost, int64_t *__restrict__ runcoff,
                                                   int64_t *__restrict__ ctotal, QueueSegs segs) {
    __shared__ int32_t wtot[256 / WAVE];
    __shared__ int64_t wtot64[256 / WAVE];
    scan_runs<int32_t>(runcnt, R, NB, runoff, counts, wtot);
    if (runcost) {
        scan_runs<int64_t>(runcost, R, NB, runcoff, ctotal, wtot64);
        // (scan_runs ends with a barrier after thread 255 wrote the totals)
        for (int k = 0; k < NB; ++k) {
            if (segs.seg[k] == nullptr) continue;
            const int64_t n = ctotal[k] == 0 ? segs.W[k] + 1 : 1;
            for (int64_t g = threadIdx.x; g < n; g += 256) segs.seg[k][g] = 0;
        }
    }
}

// Pass 1: one wave per run, from the buckets in bucket_of (and the costs).
__global__ __launch_bounds__(256) void k_queue_count(const int8_t *__restrict__ bucket_of, int64_t B, int NB,
                                                    int32_t *__restrict__ runcnt, const int32_t *__restrict__ cost_of,
                                                    int64_t *__restrict__ runcost) {
    const int64_t b0 = ((int64_t)blockIdx.x * 4 + uniform_i32((int)(threadIdx.x / WAVE))) * QRUN;
    if (b0 >= B) return;
    const int lane = lane_id();
    const int n = (int)(B - b0 < QRUN ? B - b0 : QRUN);
    const int my = lane < n ? (int)bucket_of[b0 + lane] : -1;
    const int64_t c = (cost_of && lane < n) ? (int64_t)cost_of[b0 + lane] : 0;
    const int64_t r = b0 / QRUN;
    for (int k = 0; k < NB; ++k) {
        const int cnt = __popcll(__ballot(my == k));
        if (lane == 0) runcnt[r * NB + k] = cnt;
        if (cost_of) {
            int64_t v = my == k ? c : 0;
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, WAVE);
            if (lane == 0) runcost[r * NB + k] = v;
        }
    }
}

// Pass 3: one wave per run.
__global__ __launch_bounds__(256) void k_queue_scatter(const int8_t *__restrict__ bucket_of,
                                                      const int32_t *__restrict__ runoff, int64_t B, int NB,
                                                      int32_t *__restrict__ lists, const int32_t *__restrict__ cost_of,
                                                      const int64_t *__restrict__ runcoff,
                                                      const int64_t *__restrict__ ctotal, QueueSegs segs) {
    const int64_t b0 = ((int64_t)blockIdx.x * 4 + uniform_i32((int)(threadIdx.x / WAVE))) * QRUN;
    if (b0 >= B) return;
    const int lane = lane_id();
    const int n = (int)(B - b0 < QRUN ? B - b0 : QRUN);
    const int my = lane < n ? (int)bucket_of[b0 + lane] : -1;
    const int64_t c = (cost_of && lane < n) ? (int64_t)cost_of[b0 + lane] : 0;
    const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    const int64_t r = b0 / QRUN;
    for (int k = 0; k < NB; ++k) {
        const uint64_t mask = __ballot(my == k);
        const int32_t pos = runoff[r * NB + k] + __popcll(mask & below);
        if (my == k) lists[(int64_t)k * B + pos] = (int32_t)(b0 + lane);
        if (cost_of) {
            // exclusive cost prefix of this item within the bucket list
            int64_t incl = my == k ? c : 0;
#pragma unroll
            for (int d = 1; d < WAVE; d <<= 1) {
                const int64_t up = __shfl_up(incl, d, WAVE);
                if (lane >= d) incl += up;
            }
            if (my == k) {
                const int64_t T = ctotal[k], W = segs.W[k];
                const int64_t c0 = runcoff[r * NB + k] + incl - c, c1 = c0 + c;
                const int64_t g_lo = c0 * W / T + 1, g_hi = c1 * W / T;
                for (int64_t g = g_lo; g <= g_hi; ++g) segs.seg[k][g] = pos + 1;
            }
        }
    }
}

}  // namespace

// Passes 1-3, after the caller's classification kernel filled bucket_of
// (and, with segments, cost_of = queue_cost_of(ws, B)).  With segs, every
// seg[k] must hold W[k] + 1 entries and W[k] >= 1.
inline void queue_finish(void *ws, int64_t B, int NB, int32_t *lists, int32_t *counts, hipStream_t s,
                         const QueueSegs *segs = nullptr, void *seg_ws = nullptr) {
    char *p = (char *)ws;
    const int8_t *bucket_of = (const int8_t *)p;
    const int64_t R = queue_runs(B);
    int32_t *runcnt = (int32_t *)(p + ((size_t)B + 255) / 256 * 256);
    int32_t *runoff = runcnt + (size_t)R * NB;
    const int32_t *cost_of = nullptr;
    int64_t *runcost = nullptr, *runcoff = nullptr, *ctotal = nullptr;
    QueueSegs sg{};
    if (segs && seg_ws) {
        char *q = (char *)seg_ws;
        cost_of = (const int32_t *)q;
        runcost = (int64_t *)(q + ((size_t)B * sizeof(int32_t) + 255) / 256 * 256);
        runcoff = runcost + (size_t)R * NB;
        ctotal = runcoff + (size_t)R * NB;
        sg = *segs;
    }
    hipLaunchKernelGGL(k_queue_count, dim3((unsigned)((B + 4 * QRUN - 1) / (4 * QRUN))), dim3(256), 0, s, bucket_of,
                       B, NB, runcnt, cost_of, runcost);
    hipLaunchKernelGGL(k_queue_scan, dim3(1), dim3(256), 0, s, runcnt, R, NB, runoff, counts, runcost, runcoff,
                       ctotal, sg);
    hipLaunchKernelGGL(k_queue_scatter, dim3((unsigned)((B + 4 * QRUN - 1) / (4 * QRUN))), dim3(256), 0, s,
                       bucket_of, runoff, B, NB, lists, cost_of, runcoff, ctotal, sg);
}

// The per-item buckets inside the queue workspace.
inline int8_t *queue_bucket_of(void *ws) { return (int8_t *)ws; }
// The per-item costs inside the segment workspace.
inline int32_t *queue_cost_of(void *seg_ws) { return (int32_t *)seg_ws; }

}  // namespace sr
