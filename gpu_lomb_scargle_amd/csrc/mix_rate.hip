// mix_rate: whether the FP32 VALU work of one wavefront overlaps the MFMAs of
// another on the same SIMD.
//
// Loop body: 24 independent-chain MFMAs v_mfma_f32_32x32x16_f16 (six
// accumulators, the matrix-core tile's count per 16 observations) and/or K
// FP32 FMAs on eight independent chains, no memory traffic; one or two
// wavefronts per SIMD (1024 or 2048 single-wave workgroups).  Printed: the
// time per loop iteration and wavefront, per SIMD.  If the matrix cores and the
// VALU worked side by side, two wavefronts of MFMA + VALU would take the
// larger of the two separate times; if they share the issue, the sum.
#include <hip/hip_runtime.h>
#include <cstdio>

typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));
#define CHECK(x)                                                                     \
    do {                                                                             \
        hipError_t e_ = (x);                                                         \
        if (e_ != hipSuccess) {                                                      \
            std::printf("HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__); \
            return 1;                                                                \
        }                                                                            \
    } while (0)

template <int K, bool MF>
__global__ __launch_bounds__(64) void k_mix(float *out, int iters, float seed) {
    const int l = threadIdx.x;
    half8_t a, b;
    for (int i = 0; i < 8; ++i) {
        a[i] = (_Float16)(0.37f * (float)((l * 7 + i * 13) % 29) * seed - 1.0f);
        b[i] = (_Float16)(0.11f * (float)((l * 5 + i * 3) % 31) * seed - 1.5f);
    }
    f32x16_t acc[6];
    for (int k = 0; k < 6; ++k) acc[k] = f32x16_t{};
    float x[8];
    for (int i = 0; i < 8; ++i) x[i] = seed * (float)(l + i);
    for (int it = 0; it < iters; ++it) {
        if (MF) {
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int k = 0; k < 6; ++k) acc[k] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, acc[k], 0, 0, 0);
        }
#pragma unroll
        for (int j = 0; j < K / 8; ++j)
#pragma unroll
            for (int i = 0; i < 8; ++i) x[i] = __builtin_fmaf(x[i], 0.999f, 0.001f);
        a = a * (_Float16)0.999f;
    }
    float s = 0.0f;
    for (int k = 0; k < 6; ++k)
        for (int v = 0; v < 16; ++v) s += acc[k][v];
    for (int i = 0; i < 8; ++i) s += x[i];
    out[blockIdx.x * 64 + l] = s;
}

template <int K, bool MF>
int run(float *out, int simds, hipEvent_t e0, hipEvent_t e1) {
    const int iters = 4000;
    for (int waves = 1; waves <= 2; ++waves) {
        const int blocks = simds * waves;
        hipLaunchKernelGGL((k_mix<K, MF>), dim3(blocks), dim3(64), 0, 0, out, 100, 1.0f);
        float best = 1e30f;
        for (int rep = 0; rep < 3; ++rep) {
            CHECK(hipEventRecord(e0));
            hipLaunchKernelGGL((k_mix<K, MF>), dim3(blocks), dim3(64), 0, 0, out, iters, 1.0f);
            CHECK(hipEventRecord(e1));
            CHECK(hipEventSynchronize(e1));
            float ms;
            CHECK(hipEventElapsedTime(&ms, e0, e1));
            if (ms < best) best = ms;
        }
        // time per iteration per wave-slot, in ns
        std::printf("MFMA %d, VALU FMAs %4d, waves/SIMD %d: %.3f ms, %.1f ns per iteration per SIMD\n", MF ? 24 : 0, K,
               waves, best, best * 1e6 / ((double)iters * waves));
    }
    return 0;
}

int main() {
    hipDeviceProp_t prop;
    CHECK(hipGetDeviceProperties(&prop, 0));
    const int simds = prop.multiProcessorCount * 4;
    float *out;
    CHECK(hipMalloc(&out, sizeof(float) * 64 * (size_t)simds * 2));
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    if (run<0, true>(out, simds, e0, e1) || run<256, false>(out, simds, e0, e1) ||
        run<256, true>(out, simds, e0, e1) || run<512, false>(out, simds, e0, e1) ||
        run<512, true>(out, simds, e0, e1))
        return 1;
    CHECK(hipFree(out));
    return 0;
}
