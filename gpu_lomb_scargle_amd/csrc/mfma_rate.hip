// mfma_rate: sustained throughput of v_mfma_f32_32x32x16_f16, the matrix
// instruction of the matrix-core sweep tiles, on the current GPU.
//
// Each wavefront issues back-to-back MFMAs into 4 independent 32x32 FP32
// accumulators (no memory traffic, operands in registers and changing slowly
// so that the matrix cores see live data); one or two wavefronts per SIMD
// (1024 or 2048 single-wave workgroups on 256 CUs).  The result is the dense
// FP16 rate the chip sustains under matrix load, next to its published
// 2.5 PFLOP/s; dividing by 1024 flops per SIMD-cycle gives the clock that
// rate implies.
#include <hip/hip_runtime.h>

#include <cstdio>

#define CK(x)                                                                        \
    do {                                                                             \
        hipError_t e_ = (x);                                                         \
        if (e_ != hipSuccess) {                                                      \
            std::printf("HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__); \
            return 1;                                                                \
        }                                                                            \
    } while (0)

typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));

constexpr int kAcc = 4;

__global__ __launch_bounds__(64) void k_mfma_rate(float *out, int iters, float seed) {
    const int l = threadIdx.x;
    half8_t a, b;
    for (int i = 0; i < 8; ++i) {
        a[i] = (_Float16)(0.37f * (float)((l * 7 + i * 13) % 29) * seed - 1.0f);
        b[i] = (_Float16)(0.11f * (float)((l * 5 + i * 3) % 31) * seed - 1.5f);
    }
    f32x16_t acc[kAcc];
    for (int k = 0; k < kAcc; ++k) acc[k] = f32x16_t{};
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int k = 0; k < kAcc; ++k) acc[k] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, acc[k], 0, 0, 0);
        a = a * (_Float16)0.999f;
    }
    float s = 0.0f;
    for (int k = 0; k < kAcc; ++k)
        for (int v = 0; v < 16; ++v) s += acc[k][v];
    out[blockIdx.x * 64 + l] = s;
}

int main() {
    hipDeviceProp_t prop;
    CK(hipGetDeviceProperties(&prop, 0));
    const int simds = prop.multiProcessorCount * 4;
    std::printf("device %s, %d CUs (%d SIMDs)\n", prop.gcnArchName, prop.multiProcessorCount, simds);
    float *out = nullptr;
    CK(hipMalloc(&out, sizeof(float) * 64 * (size_t)simds * 2));
    hipEvent_t e0, e1;
    CK(hipEventCreate(&e0));
    CK(hipEventCreate(&e1));
    const int iters = 20000;
    for (int waves = 1; waves <= 2; ++waves) {
        const int blocks = simds * waves;
        hipLaunchKernelGGL(k_mfma_rate, dim3(blocks), dim3(64), 0, 0, out, 100, 1.0f);  // warm-up
        CK(hipGetLastError());
        for (int rep = 0; rep < 3; ++rep) {
            CK(hipEventRecord(e0));
            hipLaunchKernelGGL(k_mfma_rate, dim3(blocks), dim3(64), 0, 0, out, iters, 1.0f);
            CK(hipGetLastError());
            CK(hipEventRecord(e1));
            CK(hipEventSynchronize(e1));
            float ms = 0.0f;
            CK(hipEventElapsedTime(&ms, e0, e1));
            const double flops = (double)blocks * iters * kAcc * 32768.0;
            const double tf = flops / (ms * 1e-3) / 1e12;
            std::printf("%d wavefront(s) per SIMD: %8.3f ms  %7.1f TFLOP/s dense FP16  (%.3f GHz at 1024 flops per "
                        "SIMD-cycle)\n", waves, ms, tf, tf * 1e12 / (1024.0 * simds) / 1e9);
        }
    }
    CK(hipFree(out));
    return 0;
}
