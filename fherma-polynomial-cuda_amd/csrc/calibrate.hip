// Throughput calibration kernels for the rooflines the multiplier is measured
// against (libpm_calibrate.so; not part of the product path):
//   pm_calibrate_mfma_i8  register-only INT8 MFMA issue rate: every wave runs
//                         independent chains of v_mfma_i32_16x16x64_i8 (or
//                         v_mfma_i32_32x32x32_i8) on operands held in registers,
//                         so no memory or LDS traffic limits it
//   pm_calibrate_mul32    32-bit integer multiply rate (v_mul_lo_u32 +
//                         v_mul_hi_u32 pairs in independent chains), the
//                         instruction mix that bounds the NTT row stage
// Both time a launch with HIP events on their own stream and return the
// elapsed milliseconds and the operation count of the launch.
#include <hip/hip_runtime.h>

#include <cstdint>

namespace {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

constexpr int kChains16 = 16;  // independent 16x16x64 accumulators per wave (64 registers)
constexpr int kChains32 = 8;   // independent 32x32x32 accumulators per wave (128 registers)

__global__ __launch_bounds__(512) void mfma16_peak(int iters, int* __restrict__ sink) {
    const int lane = threadIdx.x & 63;
    const v4i a = {lane, lane * 3 + 1, lane ^ 5, lane + 7};
    const v4i b = {lane * 5, lane + 11, lane ^ 9, lane * 7 + 2};
    v4i acc[kChains16];
#pragma unroll
    for (int j = 0; j < kChains16; ++j) acc[j] = v4i{j, 0, 0, 0};
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int j = 0; j < kChains16; ++j) acc[j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b, acc[j], 0, 0, 0);
    }
    int s = 0;
#pragma unroll
    for (int j = 0; j < kChains16; ++j) s += acc[j][0] ^ acc[j][1] ^ acc[j][2] ^ acc[j][3];
    if (s == 0x7fffffff) sink[blockIdx.x] = s;  // keeps the chains live; practically never taken
}

__global__ __launch_bounds__(512) void mfma32_peak(int iters, int* __restrict__ sink) {
    const int lane = threadIdx.x & 63;
    const v4i a = {lane, lane * 3 + 1, lane ^ 5, lane + 7};
    const v4i b = {lane * 5, lane + 11, lane ^ 9, lane * 7 + 2};
    v16i acc[kChains32];
#pragma unroll
    for (int j = 0; j < kChains32; ++j) {
        acc[j] = v16i{};
        acc[j][0] = j;
    }
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int j = 0; j < kChains32; ++j) acc[j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, acc[j], 0, 0, 0);
    }
    int s = 0;
#pragma unroll
    for (int j = 0; j < kChains32; ++j)
#pragma unroll
        for (int e = 0; e < 16; ++e) s ^= acc[j][e];
    if (s == 0x7fffffff) sink[blockIdx.x] = s;
}

__global__ __launch_bounds__(256) void mul32_peak(int iters, uint32_t k1, uint32_t k2, uint32_t* __restrict__ sink) {
    constexpr int kMulChains = 8;
    uint32_t x[kMulChains];
#pragma unroll
    for (int j = 0; j < kMulChains; ++j) x[j] = threadIdx.x * 2654435761u + j;
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int j = 0; j < kMulChains; ++j) x[j] = __umulhi(x[j], k1) ^ (x[j] * k2);
    }
    uint32_t s = 0;
#pragma unroll
    for (int j = 0; j < kMulChains; ++j) s += x[j];
    if (s == 0x12345u) sink[blockIdx.x] = s;
}

template <class Launch>
int timed(Launch&& launch, float* ms) {
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = 0;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess) {
        rc = -2;
    } else {
        launch(st);  // warm-up (code object load, clocks)
        (void)hipEventRecord(e0, st);
        launch(st);
        (void)hipEventRecord(e1, st);
        if (hipGetLastError() != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(ms, e0, e1) != hipSuccess)
            rc = -3;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    return rc;
}

}  // namespace

extern "C" {

// shape 16: v_mfma_i32_16x16x64_i8, shape 32: v_mfma_i32_32x32x32_i8.
// blocks x (64 * waves) threads, waves in 1..8; iters >= 1.  ops = 2 per MAC.
int pm_calibrate_mfma_i8(int shape, int blocks, int waves, int iters, float* ms, double* ops) {
    if ((shape != 16 && shape != 32) || blocks < 1 || waves < 1 || waves > 8 || iters < 1 || !ms || !ops) return -1;
    int* sink = nullptr;
    if (hipMalloc(&sink, size_t(blocks) * sizeof(int)) != hipSuccess) return -2;
    const int rc = timed(
        [&](hipStream_t st) {
            if (shape == 16)
                hipLaunchKernelGGL(mfma16_peak, dim3(blocks), dim3(64 * waves), 0, st, iters, sink);
            else
                hipLaunchKernelGGL(mfma32_peak, dim3(blocks), dim3(64 * waves), 0, st, iters, sink);
        },
        ms);
    (void)hipFree(sink);
    const double per_mfma = 2.0 * (shape == 16 ? 16.0 * 16 * 64 : 32.0 * 32 * 32);
    *ops = double(blocks) * waves * iters * (shape == 16 ? kChains16 : kChains32) * per_mfma;
    return rc;
}

// blocks x 256 threads; ops = 32-bit multiplies (one v_mul_lo_u32 and one
// v_mul_hi_u32 per chain step) summed over all lanes.
int pm_calibrate_mul32(int blocks, int iters, float* ms, double* ops) {
    if (blocks < 1 || iters < 1 || !ms || !ops) return -1;
    uint32_t* sink = nullptr;
    if (hipMalloc(&sink, size_t(blocks) * sizeof(uint32_t)) != hipSuccess) return -2;
    const int rc = timed(
        [&](hipStream_t st) {
            hipLaunchKernelGGL(mul32_peak, dim3(blocks), dim3(256), 0, st, iters, 0x9E3779B9u, 0x85EBCA6Bu, sink);
        },
        ms);
    (void)hipFree(sink);
    *ops = double(blocks) * 256.0 * iters * 8 * 2;
    return rc;
}

}  // extern "C"
