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
#pragma once

#include "sr_wave.h"
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sr {

constexpr int QRUN = 64;  // items per classification run (one wave)

inline int64_t queue_runs(int64_t B) { return (B + QRUN - 1) / QRUN; }
inline size_t queue_ws_bytes(int64_t B, int NB) {
    return (((size_t)B + 255) / 256 * 256) + 2 * (size_t)queue_runs(B) * NB * sizeof(int32_t) + 256;
}

namespace {

// Pass 2: one block of 256 threads; thread t owns a contiguous range of runs.
// The 256 thread totals are scanned per wave with shuffles, then across the
// 4 wave totals (integer sums: the offsets do not depend on the order).
__global__ __launch_bounds__(256) void k_queue_scan(const int32_t *__restrict__ runcnt, int64_t R, int NB,
                                                   int32_t *__restrict__ runoff, int32_t *__restrict__ counts) {
    __shared__ int32_t wtot[256 / WAVE];
    const int t = threadIdx.x, lane = lane_id(), w = t / WAVE;
    const int64_t per = (R + 255) / 256;
    const int64_t r0 = t * per, r1 = r0 + per < R ? r0 + per : R;
    for (int k = 0; k < NB; ++k) {
        int32_t s = 0;
        for (int64_t r = r0; r < r1; ++r) s += runcnt[r * NB + k];
        int32_t incl = s;
#pragma unroll
        for (int d = 1; d < WAVE; d <<= 1) {
            const int32_t up = __shfl_up(incl, d, WAVE);
            if (lane >= d) incl += up;
        }
        if (lane == WAVE - 1) wtot[w] = incl;
        __syncthreads();
        int32_t before = 0;
        for (int q = 0; q < w; ++q) before += wtot[q];
        if (t == 255) counts[k] = before + incl;
        int32_t acc = before + incl - s;
        for (int64_t r = r0; r < r1; ++r) {
            runoff[r * NB + k] = acc;
            acc += runcnt[r * NB + k];
        }
        __syncthreads();  // wtot is rewritten for the next bucket
    }
}

// Pass 1: one wave per run, from the buckets in bucket_of.
__global__ __launch_bounds__(256) void k_queue_count(const int8_t *__restrict__ bucket_of, int64_t B, int NB,
                                                    int32_t *__restrict__ runcnt) {
    const int64_t b0 = ((int64_t)blockIdx.x * 4 + uniform_i32((int)(threadIdx.x / WAVE))) * QRUN;
    if (b0 >= B) return;
    const int lane = lane_id();
    const int n = (int)(B - b0 < QRUN ? B - b0 : QRUN);
    const int my = lane < n ? (int)bucket_of[b0 + lane] : -1;
    const int64_t r = b0 / QRUN;
    for (int k = 0; k < NB; ++k) {
        const int cnt = __popcll(__ballot(my == k));
        if (lane == 0) runcnt[r * NB + k] = cnt;
    }
}

// Pass 3: one wave per run.
__global__ __launch_bounds__(256) void k_queue_scatter(const int8_t *__restrict__ bucket_of,
                                                      const int32_t *__restrict__ runoff, int64_t B, int NB,
                                                      int32_t *__restrict__ lists) {
    const int64_t b0 = ((int64_t)blockIdx.x * 4 + uniform_i32((int)(threadIdx.x / WAVE))) * QRUN;
    if (b0 >= B) return;
    const int lane = lane_id();
    const int n = (int)(B - b0 < QRUN ? B - b0 : QRUN);
    const int my = lane < n ? (int)bucket_of[b0 + lane] : -1;
    const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    const int64_t r = b0 / QRUN;
    for (int k = 0; k < NB; ++k) {
        const uint64_t mask = __ballot(my == k);
        if (my == k) lists[(int64_t)k * B + runoff[r * NB + k] + __popcll(mask & below)] = (int32_t)(b0 + lane);
    }
}

}  // namespace

// Passes 1-3, after the caller's classification kernel filled bucket_of.
inline void queue_finish(void *ws, int64_t B, int NB, int32_t *lists, int32_t *counts, hipStream_t s) {
    char *p = (char *)ws;
    const int8_t *bucket_of = (const int8_t *)p;
    const int64_t R = queue_runs(B);
    int32_t *runcnt = (int32_t *)(p + ((size_t)B + 255) / 256 * 256);
    int32_t *runoff = runcnt + (size_t)R * NB;
    hipLaunchKernelGGL(k_queue_count, dim3((unsigned)((B + 4 * QRUN - 1) / (4 * QRUN))), dim3(256), 0, s, bucket_of,
                       B, NB, runcnt);
    hipLaunchKernelGGL(k_queue_scan, dim3(1), dim3(256), 0, s, runcnt, R, NB, runoff, counts);
    hipLaunchKernelGGL(k_queue_scatter, dim3((unsigned)((B + 4 * QRUN - 1) / (4 * QRUN))), dim3(256), 0, s,
                       bucket_of, runoff, B, NB, lists);
}

// The per-item buckets inside the queue workspace.
inline int8_t *queue_bucket_of(void *ws) { return (int8_t *)ws; }

}  // namespace sr
