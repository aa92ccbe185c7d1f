// Measurement probe: VALU issue rate of one FP32 operation kind on gfx950.
//
// Every thread runs 8 independent chains of `iters` operations of one kind
// (FMA, or one of the hardware transcendentals v_exp_f32, v_log_f32,
// v_sin_f32, v_rcp_f32, v_sqrt_f32, v_rsq_f32), so that the rate is set by
// issue throughput, not latency.  The ratio of the FMA rate to a
// transcendental's rate is the cost of that transcendental in FMA slots,
// which the benchmark's algorithmic operation count uses.  The chains'
// values are not meaningful (they may reach inf or NaN, which does not
// change the issue rate); their sum is written so the work is kept.
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace {

template <int KIND>
__device__ inline float step(float x) {
    if constexpr (KIND == 0) return __builtin_fmaf(x, 0.9999f, 1e-4f);
    else if constexpr (KIND == 1) return __builtin_amdgcn_exp2f(-x);
    else if constexpr (KIND == 2) return __builtin_amdgcn_logf(x);
    else if constexpr (KIND == 3) return __builtin_amdgcn_sinf(x);
    else if constexpr (KIND == 4) return __builtin_amdgcn_rcpf(x);
    else if constexpr (KIND == 5) return __builtin_amdgcn_sqrtf(x);
    else return __builtin_amdgcn_rsqf(x);
}

template <int KIND>
__global__ __launch_bounds__(256) void k_probe(int iters, float *out) {
    float a[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) a[j] = 0.25f + 0.01f * (float)(threadIdx.x + j);
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int j = 0; j < 8; ++j) a[j] = step<KIND>(a[j]);
    }
    float s = 0.f;
#pragma unroll
    for (int j = 0; j < 8; ++j) s += a[j];
    out[(int64_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

template <int KIND>
int launch(int64_t blocks, int iters, float *out, hipStream_t s) {
    hipLaunchKernelGGL(k_probe<KIND>, dim3((unsigned)blocks), dim3(256), 0, s, iters, out);
    return (int)hipGetLastError();
}

}  // namespace

extern "C" {

int sr_probe_valu_ops_per_thread(int iters) { return iters < 0 ? -1 : 8 * iters; }

int sr_probe_valu_rate(int kind, int64_t blocks, int iters, float *out, void *stream) {
    if (blocks <= 0 || blocks > 0x7fffffff || iters <= 0 || out == nullptr) return -1;
    hipStream_t s = (hipStream_t)stream;
    switch (kind) {
        case 0: return launch<0>(blocks, iters, out, s);
        case 1: return launch<1>(blocks, iters, out, s);
        case 2: return launch<2>(blocks, iters, out, s);
        case 3: return launch<3>(blocks, iters, out, s);
        case 4: return launch<4>(blocks, iters, out, s);
        case 5: return launch<5>(blocks, iters, out, s);
        case 6: return launch<6>(blocks, iters, out, s);
        default: return -1;
    }
}

}  // extern "C"
