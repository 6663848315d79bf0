// valu_rates: issue rates of the VALU operations the Lomb-Scargle sweep is
// built from, and the accuracy of the hardware sin/cos, on the current GPU.
//
// The sweep's roofline counts work in VALU lane-cycles per frequency x
// observation term; this tool measures what one lane-cycle buys on the
// device: FP32/FP64 FMA, FP32 multiply, v_sin_f32 / v_cos_f32 (input in
// revolutions), and a sin + 4 FMA mix, each with 16-32 independent chains per
// thread and 32 waves per CU so that issue, not latency, is the limit.
// Output: lane-instructions per CU-cycle at the device's peak clock (for the
// sweep-body kernel: bins per CU-cycle; its instruction count comes from the
// SQ_INSTS_VALU counter).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <vector>

#define CK(x)                                                                        \
    do {                                                                             \
        hipError_t e_ = (x);                                                         \
        if (e_ != hipSuccess) {                                                      \
            std::printf("HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__); \
            return 1;                                                                \
        }                                                                            \
    } while (0)

template <int OP>
__global__ __launch_bounds__(256) void k_rate(float *out, double *outd, int iters, float a) {
    float x[32];
    double d[16];
#pragma unroll
    for (int c = 0; c < 32; ++c) x[c] = 1e-3f * (threadIdx.x + c);
#pragma unroll
    for (int c = 0; c < 16; ++c) d[c] = 1e-3 * (threadIdx.x + c) + blockIdx.x;
    float u4[4], v4[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        u4[c] = 0.5f + 1e-4f * (threadIdx.x + c);
        v4[c] = 0.25f - 1e-5f * (threadIdx.x + 3 * c);
    }
    for (int i = 0; i < iters; ++i) {
        if constexpr (OP == 0) {
#pragma unroll
            for (int c = 0; c < 32; ++c) x[c] = fmaf(x[c], a, 0.25f);
        }
        if constexpr (OP == 1) {
#pragma unroll
            for (int c = 0; c < 16; ++c) d[c] = fma(d[c], 0.999, 0.25);
        }
        if constexpr (OP == 2) {
#pragma unroll
            for (int c = 0; c < 32; ++c) x[c] = __builtin_amdgcn_sinf(x[c]);
        }
        if constexpr (OP == 3) {
#pragma unroll
            for (int c = 0; c < 32; ++c) {
                const float t = __builtin_amdgcn_sinf(x[c]);
                x[c] = fmaf(fmaf(fmaf(fmaf(t, a, x[c]), a, 0.1f), a, 0.2f), a, 0.3f);
            }
        }
        if constexpr (OP == 4) {
#pragma unroll
            for (int c = 0; c < 32; ++c) x[c] = x[c] * a;
        }
        if constexpr (OP == 5) {
            // accumulator form (two-address v_fmac_f32): x[c] += u[c%4] * v[c%4]
#pragma unroll
            for (int c = 0; c < 32; ++c) x[c] = fmaf(u4[c & 3], v4[c & 3], x[c]);
        }
        if constexpr (OP == 6) {
            // 16 FP32 FMAs and 4 FP64 FMAs, independent
#pragma unroll
            for (int c = 0; c < 16; ++c) x[c] = fmaf(u4[c & 3], v4[c & 3], x[c]);
#pragma unroll
            for (int c = 0; c < 4; ++c) d[c] = fma(d[c], 0.999, 0.25);
        }
        if constexpr (OP == 7) {
            // the split32 sweep's per-observation body for 4 bins (standard):
            // FP64 phase reduction, hardware cos/sin, 5 FMA sums per bin and a
            // 4-instruction FP32 rotation between bins
            const double tj = d[i & 15] + (double)i;
            double ph = d[0] * 1e-3 * tj;
            ph -= __builtin_rint(ph);
            const float phf = (float)ph;
            float c = __builtin_amdgcn_cosf(phf), s = __builtin_amdgcn_sinf(phf);
            const float yv = u4[i & 3], rc = v4[0], rs = v4[1];
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                x[5 * m + 0] = fmaf(yv, c, x[5 * m + 0]);
                x[5 * m + 1] = fmaf(yv, s, x[5 * m + 1]);
                x[5 * m + 2] = fmaf(c, c, x[5 * m + 2]);
                x[5 * m + 3] = fmaf(s, s, x[5 * m + 3]);
                x[5 * m + 4] = fmaf(c, s, x[5 * m + 4]);
                if (m < 3) {
                    const float cn = fmaf(c, rc, -s * rs);
                    const float sn = fmaf(s, rc, c * rs);
                    c = cn;
                    s = sn;
                }
            }
        }
    }
    float s = 0.0f;
    for (int c = 0; c < 32; ++c) s += x[c];
    double sd = 0.0;
    for (int c = 0; c < 16; ++c) sd += d[c];
    out[blockIdx.x * 256 + threadIdx.x] = s;
    outd[blockIdx.x * 256 + threadIdx.x] = sd;
}

__global__ void k_accuracy(const float *x, float *s, float *c, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        s[i] = __builtin_amdgcn_sinf(x[i]);
        c[i] = __builtin_amdgcn_cosf(x[i]);
    }
}

template <int OP>
int rate(const char *name, double ops_per_iter, int blocks, float *o, double *od, double clock_hz,
         int cus) {
    const int iters = 4096;
    hipEvent_t e0, e1;
    CK(hipEventCreate(&e0));
    CK(hipEventCreate(&e1));
    hipLaunchKernelGGL(k_rate<OP>, dim3(blocks), dim3(256), 0, 0, o, od, 64, 0.999f);
    CK(hipDeviceSynchronize());
    CK(hipEventRecord(e0));
    hipLaunchKernelGGL(k_rate<OP>, dim3(blocks), dim3(256), 0, 0, o, od, iters, 0.999f);
    CK(hipEventRecord(e1));
    CK(hipEventSynchronize(e1));
    float ms = 0.0f;
    CK(hipEventElapsedTime(&ms, e0, e1));
    const double lane_ops = ops_per_iter * iters * (double)blocks * 256;
    std::printf("%-28s %8.3f ms  %9.1f Glane-instr/s  %6.2f lane-instr per CU-cycle\n", name, ms,
                lane_ops / ms / 1e6, lane_ops / (ms * 1e-3) / clock_hz / cus);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return 0;
}

int main() {
    int dev = 0;
    CK(hipGetDevice(&dev));
    hipDeviceProp_t p;
    CK(hipGetDeviceProperties(&p, dev));
    const int cus = p.multiProcessorCount;
    const double clock_hz = p.clockRate * 1e3;
    std::printf("device %s, %d CUs, peak clock %.0f MHz\n", p.gcnArchName, cus, clock_hz / 1e6);
    const int blocks = cus * 8;
    float *o;
    double *od;
    CK(hipMalloc(&o, (size_t)blocks * 256 * 4));
    CK(hipMalloc(&od, (size_t)blocks * 256 * 8));
    if (rate<4>("v_mul_f32 (32 chains)", 32, blocks, o, od, clock_hz, cus)) return 1;
    if (rate<0>("v_fma_f32 (32 chains)", 32, blocks, o, od, clock_hz, cus)) return 1;
    if (rate<1>("v_fma_f64 (16 chains)", 16, blocks, o, od, clock_hz, cus)) return 1;
    if (rate<2>("v_sin_f32 (32 chains)", 32, blocks, o, od, clock_hz, cus)) return 1;
    if (rate<3>("v_sin_f32 + 4 v_fma_f32", 160, blocks, o, od, clock_hz, cus)) return 1;
    if (rate<5>("v_fmac_f32 (32 accumulators)", 32, blocks, o, od, clock_hz, cus)) return 1;
    if (rate<6>("16 v_fmac_f32 + 4 v_fma_f64", 20, blocks, o, od, clock_hz, cus)) return 1;
    // sweep body: 2 trans + 20 FMA sums + 12 rotation ops + ~8 phase/convert ops;
    // counted as 1 term = 1/4 of the body, rate in lane-instructions is from the counters
    if (rate<7>("sweep body, 4 bins (x 1/4)", 1, blocks, o, od, clock_hz, cus)) return 1;

    // accuracy of v_sin_f32 / v_cos_f32 on [-0.5, 0.5) revolutions against FP64
    const int n = 1 << 22;
    std::vector<float> hx(n), hs(n), hc(n);
    for (int i = 0; i < n; ++i) hx[i] = (float)(-0.5 + (double)i / n);
    float *dx, *ds, *dc;
    CK(hipMalloc(&dx, (size_t)n * 4));
    CK(hipMalloc(&ds, (size_t)n * 4));
    CK(hipMalloc(&dc, (size_t)n * 4));
    CK(hipMemcpy(dx, hx.data(), (size_t)n * 4, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_accuracy, dim3(n / 256), dim3(256), 0, 0, dx, ds, dc, n);
    CK(hipDeviceSynchronize());
    CK(hipMemcpy(hs.data(), ds, (size_t)n * 4, hipMemcpyDeviceToHost));
    CK(hipMemcpy(hc.data(), dc, (size_t)n * 4, hipMemcpyDeviceToHost));
    double es = 0, ec = 0, rs = 0;
    for (int i = 0; i < n; ++i) {
        const double ts = std::sin(2.0 * M_PI * (double)hx[i]);
        const double tc = std::cos(2.0 * M_PI * (double)hx[i]);
        es = std::fmax(es, std::fabs(hs[i] - ts));
        ec = std::fmax(ec, std::fabs(hc[i] - tc));
        rs += (hs[i] - ts) * (hs[i] - ts);
    }
    std::printf("v_sin_f32 max abs error %.3e (rms %.3e), v_cos_f32 max abs error %.3e\n", es,
                std::sqrt(rs / n), ec);
    return 0;
}
