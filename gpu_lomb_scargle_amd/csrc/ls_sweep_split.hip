// Split-form sweep kernels (split: FP64 rotation and sums; split32: FP32
// rotation and FP32 partial sums folded into FP64, and the matrix-core
// tiles), the observation-slice combiner, and their tile registry entries.

#include "ls_device.h"

namespace lsamd {

namespace {

// ---------------------------------------------------------------------------
// Split forms: the powers of a thread's M bins from the object's FP64 sums
// (tau-free power, normalisation, periodogram store, the chunk's peaks).
//   standard:    XC, XS, CC, CS; SS = N - CC
//   generalized: C, S, YC, YS, CC, [SS,] CS over the normalised weights; SS
//                summed (SS_SUM, split32) or 1 - CC (split, whose unit-modulus
//                seeds make the weights sum to 1 to FP64 rounding)
// ---------------------------------------------------------------------------
template <bool SS_SUM, bool GEN, typename OUT, int T, int M, int NS>
__device__ __forceinline__ void split_epilogue(const SweepArgs &a, const double (&acc)[NS][M],
                                               uint32_t obj, uint32_t chunk,
                                               double (*s_pp)[kPeaks], uint32_t (*s_pi)[kPeaks]) {
    const uint32_t tid = threadIdx.x;
    const uint32_t kbase = chunk * (uint32_t)(T * M);
    const double scale = a.obj_scale[obj];
    const double nobs = (double)(a.offsets[obj + 1] - a.offsets[obj]);
    double Wf = 1.0, Yb = 0.0, YY = 1.0;
    if constexpr (GEN) {
        Wf = a.obj_gen[kGen * obj + 0];
        Yb = a.obj_gen[kGen * obj + 1];
        YY = a.obj_gen[kGen * obj + 2];
    }
    const double iW = 1.0 / Wf;  // one division per work item, not seven per bin
    OUT *pg = static_cast<OUT *>(a.pgram);
    Peaks L;
    peaks_init(L);
#pragma unroll
    for (int m = 0; m < M; ++m) {
        const uint32_t kf = kbase + m * T + tid;
        double P;
        if constexpr (GEN) {
            const double C = acc[0][m] * iW, S = acc[1][m] * iW;
            const double YC = acc[2][m] * iW - Yb * C, YS = acc[3][m] * iW - Yb * S;
            const double CCw = acc[4][m] * iW;
            const double CC = CCw - C * C;
            double SS, CS;
            if constexpr (SS_SUM) {
                SS = acc[5][m] * iW - S * S;
                CS = acc[6][m] * iW - C * S;
            } else {
                SS = (1.0 - CCw) - S * S;
                CS = acc[5][m] * iW - C * S;
            }
            const double D = CC * SS - CS * CS;
            P = (SS * YC * YC - 2.0 * CS * YC * YS + CC * YS * YS) / (D * YY);
        } else {
            const double XC = acc[0][m], XS = acc[1][m];
            const double CC = acc[2][m], CS = acc[3][m], SS = nobs - CC;
            const double D = CC * SS - CS * CS;
            P = 0.5 * (SS * XC * XC - 2.0 * CS * XC * XS + CC * XS * XS) / D;
        }
        const OUT pw = (OUT)(P * scale);
        if (kf < a.nf) {
            if (pg) pg[(uint64_t)obj * a.nf + kf] = pw;
            peaks_add(L, a.top_k, (double)pw, kf);
        }
    }
    if (a.top_k) write_partial<OUT, T>(a, L, obj, chunk, s_pp, s_pi);
}

// One observation slice's sums of a thread's M bins, for k_combine_slices.
template <int NS, int M, int T>
__device__ __forceinline__ void store_slice(const SweepArgs &a, const double (&acc)[NS][M],
                                            uint32_t obj, uint32_t chunk, uint32_t slice) {
    double *dst = a.part_sums +
                  (((uint64_t)obj * a.chunks + chunk) * a.slices + slice) * (uint64_t)(NS * M * T) +
                  threadIdx.x;
#pragma unroll
    for (int q = 0; q < NS; ++q)
#pragma unroll
        for (int m = 0; m < M; ++m) dst[(q * M + m) * T] = acc[q][m];
}

// Sums the observation slices of one (object, chunk) in slice order and
// finishes it like the single-slice sweep.  One workgroup per (object, chunk).
template <bool SS_SUM, bool GEN, typename OUT, int T, int M>
__global__ __launch_bounds__(T) void k_combine_slices(SweepArgs a) {
    constexpr int NS = GEN ? (SS_SUM ? 7 : 6) : 4;
    __shared__ double s_pp[T / 64][kPeaks];
    __shared__ uint32_t s_pi[T / 64][kPeaks];
    const uint64_t oc = blockIdx.x;
    const uint32_t obj = (uint32_t)(oc / a.chunks);
    const uint32_t chunk = (uint32_t)(oc - (uint64_t)obj * a.chunks);
    const double *src = a.part_sums + oc * a.slices * (uint64_t)(NS * M * T) + threadIdx.x;
    double acc[NS][M];
#pragma unroll
    for (int q = 0; q < NS; ++q)
#pragma unroll
        for (int m = 0; m < M; ++m) acc[q][m] = 0.0;
    for (uint32_t sl = 0; sl < a.slices; ++sl) {
#pragma unroll
        for (int q = 0; q < NS; ++q)
#pragma unroll
            for (int m = 0; m < M; ++m) acc[q][m] += src[(q * M + m) * T];
        src += NS * M * T;
    }
    split_epilogue<SS_SUM, GEN, OUT, T, M, NS>(a, acc, obj, chunk, s_pp, s_pi);
}

// ---------------------------------------------------------------------------
// Split form, FP32 sums ("split32").  Per observation and thread: one FP64
// multiply and round for the seed phase (revolutions) of the thread's middle
// bin, one hardware sin/cos pair, then the other M-1 bins in FP32 from the
// observation's precomputed step D (cos, sin of T bins); FP32 partial sums
// over 32 (standard) or 64 (generalized) observations are folded into FP64
// accumulators.
//
// The standard sweep takes the other bins from the seed in one of two ways,
// chosen per observation (so uniformly across the workgroup):
//   * |sin D| >= kRecMinSin: one rotation to the seed's upper neighbour, then
//     the three-term recurrence z[m+1] = 2 cos D z[m] - z[m-1] up and down,
//     one FMA per component and bin.  A rounding error of 2 cos D moves the
//     recurrence's angle by that error / (2 sin D), hence the bound on sin D
//     (about 84% of observations qualify).
//   * otherwise: the rotation up and down (two multiplies and two FMAs per
//     bin), whose error does not depend on D.
// Staging puts each block's recurrence observations first, so both kinds run
// as branch-free loops.  Each recurrence step is one v_fmac_f32 that writes
// over the older of its two inputs: new = X latest + older with
// X = -2 cos D g_old g_latest, which yields -g_old z[new], so the register
// holds the phasor up to a sign fixed at compile time.  The quadratic sums do
// not see that sign; the linear ones multiply by the negated weight instead.
// The generalized sweep (fewer bins per thread) always rotates: there the
// recurrence saved less than the partitioned staging costs.
// Every FP32 instruction of the inner loop is an all-VGPR v_fmac_f32 /
// v_mul_f32 (the staged records carry -sin D, -2 cos D and the negated
// weights, so that no negated operand forces the slower VOP3 encoding).
// ---------------------------------------------------------------------------
constexpr float kRecMinSin = 0.25f;

template <bool GEN, typename OUT, int T, int M>
__global__ __launch_bounds__(T) void k_sweep_split32(SweepArgs a) {
    // standard: XC, XS, CC, CS (SS = N - CC); generalized: all seven sums
    // (its centred SS - S^2 cancels, so SS is accumulated, not derived)
    constexpr int NS = GEN ? 7 : 4;
    constexpr int kSeed = (M - 1) / 2;  // the bin whose phasor is evaluated directly
    constexpr uint32_t kF = GEN ? kFlushGen : kFlushStd;
    constexpr int kWaves = T / 64;
    constexpr bool kRec = !GEN;  // recurrence class in use
    __shared__ double s_t[T];
    __shared__ float4 s_a[T];
    __shared__ float4 s_k[kRec ? T : 1];  // {2 cos D, -2 cos D, -r.x, -r.y}
    __shared__ float s_n[GEN ? T : 1];
    __shared__ uint32_t s_cnt[2][kRec ? kWaves : 1];
    __shared__ double s_pp[T / 64][kPeaks];
    __shared__ uint32_t s_pi[T / 64][kPeaks];
    __shared__ uint64_t s_item;
    uint32_t staged = ~0u;  // object whose records the LDS holds whole (at most T of them)
    uint32_t na = 0;        // recurrence observations at the head of the staged block
    for (uint64_t item = first_item(a, &s_item); item < a.total_items;
         item = next_item(a, &s_item, item)) {
        uint32_t obj, chunk, slice;
        uint64_t beg, end;
        item_coords_sliced(a, item, obj, chunk, slice, beg, end);
        const uint32_t tid = threadIdx.x;
        const uint32_t kbase = chunk * (uint32_t)(T * M);
        const double fseed = a.fmin_rev + (double)(kbase + kSeed * T + tid) * a.df_rev;

        double acc[NS][M];
    #pragma unroll
        for (int q = 0; q < NS; ++q)
    #pragma unroll
            for (int m = 0; m < M; ++m) acc[q][m] = 0.0;

        // the next chunk of an object of at most T observations reuses them
        const bool whole = a.slices == 1 && end - beg <= (uint64_t)T;
        const bool reuse = whole && obj == staged;
        staged = whole ? obj : ~0u;
        for (uint64_t b0 = beg; b0 < end; b0 += T) {
            const uint32_t cnt = (uint32_t)((end - b0) < (uint64_t)T ? (end - b0) : (uint64_t)T);
            __syncthreads();  // also orders the previous item's peak merge before this one's
            if (!reuse && !kRec) {
                if (tid < cnt) {
                    s_t[tid] = a.rec_t[b0 + tid];
                    s_a[tid] = a.rec_a[b0 + tid];
                    if constexpr (GEN) s_n[tid] = a.rec_n[b0 + tid];
                }
                __syncthreads();
            }
            if constexpr (kRec) {
                if (!reuse) {
                    // stage the block with its recurrence observations first
                    // (each kind in index order)
                    const bool in = tid < cnt;
                    double rt = 0.0;
                    float4 ra = make_float4(0.f, 0.f, 0.f, 0.f);
                    if (in) {
                        rt = a.rec_t[b0 + tid];
                        ra = a.rec_a[b0 + tid];
                    }
                    const bool rec = in && fabsf(ra.w) >= kRecMinSin;
                    const uint64_t m_rec = __ballot(rec);
                    const uint64_t m_rot = __ballot(in) & ~m_rec;
                    const uint32_t lane = tid & 63u, wave = tid >> 6;
                    const uint64_t below = (1ull << lane) - 1ull;
                    if (lane == 0) {
                        s_cnt[0][wave] = (uint32_t)__popcll(m_rec);
                        s_cnt[1][wave] = (uint32_t)__popcll(m_rot);
                    }
                    __syncthreads();
                    uint32_t pre_rec = 0, pre_rot = 0, n_rec = 0;
    #pragma unroll
                    for (int w = 0; w < kWaves; ++w) {
                        const uint32_t cr = s_cnt[0][w], co = s_cnt[1][w];
                        if ((uint32_t)w < wave) {
                            pre_rec += cr;
                            pre_rot += co;
                        }
                        n_rec += cr;
                    }
                    if (in) {
                        const uint32_t dst =
                            rec ? pre_rec + (uint32_t)__popcll(m_rec & below)
                                : n_rec + pre_rot + (uint32_t)__popcll(m_rot & below);
                        const float k2 = ra.z + ra.z;
                        s_t[dst] = rt;
                        s_a[dst] = ra;
                        s_k[dst] = make_float4(k2, -k2, -ra.x, -ra.y);
                    }
                    na = n_rec;
                    __syncthreads();
                }
            }
            for (uint32_t j0 = 0; j0 < cnt; j0 += kF) {
                const uint32_t j1 = (j0 + kF < cnt) ? j0 + kF : cnt;
                // recurrence observations in pairs (a lone one takes the
                // rotation loop, which handles any observation)
                uint32_t jm = na < j0 ? j0 : (na > j1 ? j1 : na);
                jm = j0 + ((jm - j0) & ~1u);
                float p[NS][M];
    #pragma unroll
                for (int q = 0; q < NS; ++q)
    #pragma unroll
                    for (int m = 0; m < M; ++m) p[q][m] = 0.0f;
                // one observation: [j0, jm) by the recurrence, [jm, j1) by rotation
                auto body = [&](uint32_t j, auto recur) {
                    const double tj = s_t[j];
                    const float4 r = s_a[j];
                    const float nsd = GEN ? s_n[j] : r.y;  // -sin D
                    double ph = fseed * tj;
                    ph -= __builtin_rint(ph);
                    const float phf = (float)ph;
                    float c = __builtin_amdgcn_cosf(phf);
                    float s = __builtin_amdgcn_sinf(phf);
                    if constexpr (GEN) {
                        // u = sqrt(w) z: the steps are linear, and every
                        // weighted sum becomes one FMA (w c = sqrt(w) u_c,
                        // w c^2 = u_c^2, w y c = (sqrt(w) y) u_c)
                        c *= r.y;
                        s *= r.y;
                    }
                    // bin m's sums from a register pair holding g z[m]
                    auto add = [&](int m, float cv, float sv, float wy, float wx) {
                        if constexpr (GEN) {
                            p[0][m] = fmaf(wy, cv, p[0][m]);
                            p[1][m] = fmaf(wy, sv, p[1][m]);
                            p[2][m] = fmaf(wx, cv, p[2][m]);
                            p[3][m] = fmaf(wx, sv, p[3][m]);
                            p[4][m] = fmaf(cv, cv, p[4][m]);
                            p[5][m] = fmaf(sv, sv, p[5][m]);
                            p[6][m] = fmaf(cv, sv, p[6][m]);
                        } else {
                            p[0][m] = fmaf(wx, cv, p[0][m]);
                            p[1][m] = fmaf(wx, sv, p[1][m]);
                            p[2][m] = fmaf(cv, cv, p[2][m]);
                            p[3][m] = fmaf(cv, sv, p[3][m]);
                        }
                    };
                    // The seed sits at bin kSeed; the steps walk up to bin
                    // M-1 and down to bin 0, so no bin is more than M/2
                    // steps from its seed.
                    add(kSeed, c, s, r.y, r.x);
                    const float cu = fmaf(c, r.z, s * nsd);  // bin kSeed + 1
                    const float su = fmaf(s, r.z, c * r.w);
                    if constexpr (kSeed + 1 < M) add(kSeed + 1, cu, su, r.y, r.x);
                    if constexpr (kRec && decltype(recur)::value) {
                        const float4 k = s_k[j];
                        // (older, latest) pairs and their signs; each step
                        // overwrites the older with X latest + older
                        auto walk = [&](float c0, float s0, float c1, float s1, int m0, int dm,
                                        int nsteps) {
                            int g0 = 1, g1 = 1;
    #pragma unroll
                            for (int i = 0; i < nsteps; ++i) {
                                const float X = (g0 * g1 > 0) ? k.y : k.x;
                                c0 = fmaf(X, c1, c0);
                                s0 = fmaf(X, s1, s0);
                                const int gn = -g0;
                                const float ct = c1, st = s1;
                                c1 = c0;
                                s1 = s0;
                                c0 = ct;
                                s0 = st;
                                g0 = g1;
                                g1 = gn;
                                if (gn > 0)
                                    add(m0 + i * dm, c1, s1, r.y, r.x);
                                else
                                    add(m0 + i * dm, c1, s1, k.w, k.z);
                            }
                        };
                        walk(c, s, cu, su, kSeed + 2, 1, M - kSeed - 2);
                        walk(cu, su, c, s, kSeed - 1, -1, kSeed);
                    } else {
                        float cv = cu, sv = su;
    #pragma unroll
                        for (int m = kSeed + 2; m < M; ++m) {
                            const float cn = fmaf(cv, r.z, sv * nsd);
                            const float sn = fmaf(sv, r.z, cv * r.w);
                            cv = cn;
                            sv = sn;
                            add(m, cv, sv, r.y, r.x);
                        }
    #pragma unroll
                        for (int m = kSeed - 1; m >= 0; --m) {
                            const float cn = fmaf(c, r.z, s * r.w);
                            const float sn = fmaf(s, r.z, c * nsd);
                            c = cn;
                            s = sn;
                            add(m, c, s, r.y, r.x);
                        }
                    }
                };
                if constexpr (kRec) {
                    for (uint32_t j = j0; j < jm; j += 2) {
                        body(j, std::true_type{});
                        body(j + 1, std::true_type{});
                    }
                }
    #pragma unroll 2
                for (uint32_t j = jm; j < j1; ++j) body(j, std::false_type{});
    #pragma unroll
                for (int q = 0; q < NS; ++q)
    #pragma unroll
                    for (int m = 0; m < M; ++m) acc[q][m] += (double)p[q][m];
            }
        }

        if (a.slices > 1) {
            store_slice<NS, M, T>(a, acc, obj, chunk, slice);
            continue;
        }
        split_epilogue<true, GEN, OUT, T, M, NS>(a, acc, obj, chunk, s_pp, s_pi);
    }
}

// ---------------------------------------------------------------------------
// Split form, FP64 recurrence and sums ("split", the accurate fast form).  The
// seed of each (thread, observation) is the hardware FP32 sin/cos of the
// FP64-reduced phase, corrected to first order by the FP64 residual of that
// phase's rounding to FP32; the M bins then follow in FP64 from the
// observation's FP64 step (one rotation, then the three-term recurrence) and
// accumulate directly in FP64.  This removes the FP32 summation error that
// dominates at ill-conditioned (aliasing) frequencies.  The seed is brought
// to unit modulus by one Newton step, so sum c^2 + sum s^2 equals the sum of
// the weights to FP64 rounding and SS is not accumulated: SS = N - CC
// (standard) or W - CC (generalized, normalised weights).
// ---------------------------------------------------------------------------
template <bool GEN, typename OUT, int T, int M>
__global__ __launch_bounds__(T) void k_sweep_split64(SweepArgs a) {
    constexpr int NS = GEN ? 6 : 4;
    __shared__ double s_t[T];
    __shared__ double2 s_r[T];
    __shared__ float2 s_a[T];
    __shared__ double s_pp[T / 64][kPeaks];
    __shared__ uint32_t s_pi[T / 64][kPeaks];
    __shared__ uint64_t s_item;
    uint32_t staged = ~0u;  // object whose records the LDS holds whole (at most T of them)
    for (uint64_t item = first_item(a, &s_item); item < a.total_items;
         item = next_item(a, &s_item, item)) {
        uint32_t obj, chunk, slice;
        uint64_t beg, end;
        item_coords_sliced(a, item, obj, chunk, slice, beg, end);
        const uint32_t tid = threadIdx.x;
        const uint32_t kbase = chunk * (uint32_t)(T * M);
        const double fseed = a.fmin_rev + (double)(kbase + tid) * a.df_rev;

        double acc[NS][M];
    #pragma unroll
        for (int q = 0; q < NS; ++q)
    #pragma unroll
            for (int m = 0; m < M; ++m) acc[q][m] = 0.0;

        // the next chunk of an object of at most T observations reuses them
        const bool whole = a.slices == 1 && end - beg <= (uint64_t)T;
        const bool reuse = whole && obj == staged;
        staged = whole ? obj : ~0u;
        for (uint64_t b0 = beg; b0 < end; b0 += T) {
            const uint32_t cnt = (uint32_t)((end - b0) < (uint64_t)T ? (end - b0) : (uint64_t)T);
            __syncthreads();  // also orders the previous item's peak merge before this one's
            if (!reuse) {
                if (tid < cnt) {
                    s_t[tid] = a.rec_t[b0 + tid];
                    s_r[tid] = a.rec_r[b0 + tid];
                    const float4 r4 = a.rec_a[b0 + tid];
                    s_a[tid] = make_float2(r4.x, r4.y);
                }
                __syncthreads();
            }
            for (uint32_t j = 0; j < cnt; ++j) {
                const double tj = s_t[j];
                const double2 st = s_r[j];
                const float2 yw = s_a[j];
                double ph = fseed * tj;
                ph -= __builtin_rint(ph);
                const float phf = (float)ph;
                const float cf = __builtin_amdgcn_cosf(phf);
                const float sf = __builtin_amdgcn_sinf(phf);
                const double res = (ph - (double)phf) * kTwoPi;  // exact residual, radians
                double c = fma(-res, (double)sf, (double)cf);
                double s = fma(res, (double)cf, (double)sf);
                // one Newton step towards |z| = 1: the hardware sin/cos leave
                // |z|^2 - 1 ~ 1e-7, which SS = W - CC would otherwise carry
                double nf = fma(-0.5, fma(c, c, s * s), 1.5);
                const double yv = (double)yw.x;
                const double qv = (double)yw.y;  // generalized: sqrt(w)
                if constexpr (GEN) nf *= qv;     // u = sqrt(w) z, |u|^2 = w
                c *= nf;
                s *= nf;
                auto accumulate = [&](int m, double cv, double sv) {
                    if constexpr (GEN) {
                        acc[0][m] = fma(qv, cv, acc[0][m]);
                        acc[1][m] = fma(qv, sv, acc[1][m]);
                        acc[2][m] = fma(yv, cv, acc[2][m]);
                        acc[3][m] = fma(yv, sv, acc[3][m]);
                        acc[4][m] = fma(cv, cv, acc[4][m]);
                        acc[5][m] = fma(cv, sv, acc[5][m]);
                    } else {
                        acc[0][m] = fma(yv, cv, acc[0][m]);
                        acc[1][m] = fma(yv, sv, acc[1][m]);
                        acc[2][m] = fma(cv, cv, acc[2][m]);
                        acc[3][m] = fma(cv, sv, acc[3][m]);
                    }
                };
                // Bin 1 is bin 0 rotated by the observation's step D; bins m >= 2
                // follow from the three-term recurrence z[m+1] = 2 cos(D) z[m] -
                // z[m-1] (2 FMA per bin instead of 4).  In FP64 its error stays
                // below 2e-14 over 16 bins for every D, including D near 0 and pi
                // (the rounding of cos D only perturbs the angle by O(D^2) there).
                const double twocd = 2.0 * st.x;
                double cp = c, sp = s;
                accumulate(0, c, s);
                if constexpr (M > 1) {
                    c = fma(cp, st.x, -sp * st.y);
                    s = fma(sp, st.x, cp * st.y);
                    accumulate(1, c, s);
                }
    #pragma unroll
                for (int m = 2; m < M; ++m) {
                    const double cn = fma(twocd, c, -cp);
                    const double sn = fma(twocd, s, -sp);
                    cp = c;
                    sp = s;
                    c = cn;
                    s = sn;
                    accumulate(m, c, s);
                }
            }
        }

        if (a.slices > 1) {
            store_slice<NS, M, T>(a, acc, obj, chunk, slice);
            continue;
        }
        split_epilogue<false, GEN, OUT, T, M, NS>(a, acc, obj, chunk, s_pp, s_pi);
    }
}


// ---------------------------------------------------------------------------
// Matrix-core tile of split32 (standard variant here, generalized in
// k_sweep_mfma_gen).  A work item is a block of
// 32 x 32 bins, bin = chunk * 1024 + 32 a + b.  Its sums are complex products
// of a row factor (the frequency f_a of column 0) and a column factor (b df):
//   Z1[a][b] = sum_j (y_j e^{2 pi i f_a t_j}) e^{2 pi i b df t_j} = XC + i XS
//   Z2[a][b] = sum_j e^{4 pi i f_a t_j} e^{4 pi i b df t_j}          = C2 + i S2
// with CC = (N + C2) / 2, CS = S2 / 2, SS = N - CC.  Each complex sum is four
// real products over blocks of 16 observations, formed by
// v_mfma_f32_32x32x16_f16 with FP32 accumulation.  Every factor, scaled by
// 2^10, is split into an FP16 head h and a tail l = fp16(x - h); head x head
// and the two cross products carry ~22 bits (the dropped tail x tail term is
// 2^-22 of the product); each complex sum's head x head and cross products
// accumulate in FP32 (apart, or in one sum for Z2 of the standard variant),
// unscaled by 2^-20 in FP64.  The 2^10 scale keeps FP16 subnormal heads (which the VALU
// conversions and the matrix core do not treat alike) to |x| < 2^-24; a
// subnormal tail costs at most 2^-24 absolute.  Phases are reduced in FP64
// and evaluated by the hardware sin/cos, as in the VALU split32 sweep.  The
// factors of Z2 are divided by their squared modulus (hardware sin/cos leave
// |z|^2 - 1 ~ 1e-7, and CC = (N + C2) / 2 assumes unit modulus).
// The column factor depends only on the object: k_colfac writes its operand
// fragments once per 16 observations, and y is scaled per object by a power
// of two to max |y| in [0.5, 1) (FP16 range; undone exactly in the power).
// Operand layout of 32x32x16 (checked on the device): lane l holds row (A) or
// column (B) l % 32 and observations 8 (l / 32) .. +7 of the block; the
// result holds column l % 32, rows 8 (v / 4) + 4 (l / 32) + v % 4, v < 16.
// ---------------------------------------------------------------------------
typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));

constexpr float kFacScale = 1024.0f;  // 2^10 on each factor

__device__ __forceinline__ void split16(float x, _Float16 &h, _Float16 &l) {
    h = (_Float16)(fabsf(x) < 0x1p-14f ? 0.0f : x);  // no FP16 subnormal heads
    l = (_Float16)(x - (float)h);                     // x - h is exact in FP32
}

// First column-factor slot of an object: the slots of object k are
// [floor(beg/16) + k, floor(beg/16) + k + ceil(N/16)), which never reach the
// next object's floor((beg + N)/16) + k + 1.
__device__ __forceinline__ uint64_t colfac_slot(uint64_t beg, uint32_t obj) { return (beg >> 4) + obj; }

// Wavefronts (object, y) with y < gridDim.y: the y scale (y = 0), then
// kColfacFrags fragments per 16 observations for blocks y, y + gridDim.y, ...:
// v_r, v_i, -v_i and w_r, w_i, -w_i (w = v^2), head and tail.  Observations
// past the object's end have v = 0.
__global__ __launch_bounds__(64) void k_colfac(SweepArgs a) {
    const uint32_t k = blockIdx.x, l = threadIdx.x;
    const uint64_t beg = a.offsets[k], end = a.offsets[k + 1];
    const uint32_t n = (uint32_t)(end - beg);
    if (blockIdx.y == 0) {
        float m = 0.0f;
        for (uint64_t i = beg + l; i < end; i += 64) m = fmaxf(m, fabsf(a.rec_a[i].x));
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
        if (l == 0) {
            int e = 0;
            (void)frexpf(m, &e);  // m = f 2^e, f in [0.5, 1)
            a.obj_ys[k] = m > 0.0f ? ldexpf(1.0f, -e) : 1.0f;
        }
    }
    const double fb = (double)(l & 31u) * a.df_rev;
    uint4 *dst = a.colfac + colfac_slot(beg, k) * (uint64_t)(kColfacFrags * 64) + l;
    for (uint32_t q = blockIdx.y; q * 16u < n; q += gridDim.y) {
        half8_t f[kColfacFrags];
#pragma unroll
        for (int v = 0; v < 8; ++v) {
            const uint32_t j = q * 16u + 8u * (l >> 5) + (uint32_t)v;
            float cb = 0.0f, sb = 0.0f;
            if (j < n) {
                double pb = fb * a.rec_t[beg + j];
                pb -= __builtin_rint(pb);
                cb = __builtin_amdgcn_cosf((float)pb);
                sb = __builtin_amdgcn_sinf((float)pb);
            }
            _Float16 h, t;
            split16(cb * kFacScale, h, t);
            f[0][v] = h;
            f[1][v] = t;
            split16(sb * kFacScale, h, t);
            f[2][v] = h;
            f[3][v] = t;
            f[4][v] = -h;
            f[5][v] = -t;
            // w = v^2 / |v|^2 (1/n ~ 2 - n for n ~ 1)
            const float ccb = cb * cb, ssb = sb * sb, gb = kFacScale * (2.0f - (ccb + ssb));
            split16((ccb - ssb) * gb, h, t);
            f[6][v] = h;
            f[7][v] = t;
            split16(2.0f * cb * sb * gb, h, t);
            f[8][v] = h;
            f[9][v] = t;
            f[10][v] = -h;
            f[11][v] = -t;
        }
#pragma unroll
        for (int i = 0; i < kColfacFrags; ++i) {
            uint4 u;
            __builtin_memcpy(&u, &f[i], sizeof(u));
            dst[(uint64_t)(q * kColfacFrags + i) * 64] = u;
        }
    }
}

#define LS_MFMA(ACC, A, B) ACC = __builtin_amdgcn_mfma_f32_32x32x16_f16(A, B, ACC, 0, 0, 0)

// Two waves per SIMD: six FP32 accumulators (96 registers; Z1 keeps its
// cross products apart, where merging them doubles the rounding error of the
// power, while Z2 takes all three products in one) leave room for a second
// workgroup's registers, and the fold of the 256-observation partials stays
// in FP32 (16 KiB of LDS, 8 workgroups per CU).
template <typename OUT>
__global__ __launch_bounds__(64, 2) void k_sweep_mfma(SweepArgs a) {
    constexpr int T = 64, M = 16;  // 64 threads, 1024 bins per item
    constexpr uint32_t kFoldObs = 256;  // MFMA accumulation length
    __shared__ double s_t[T];
    __shared__ float s_y[T];
    __shared__ float s_acc[4][16][T];  // sums of the 256-observation partials XC, XS, C2, S2
    __shared__ double s_pp[1][kPeaks];
    __shared__ uint32_t s_pi[1][kPeaks];
    __shared__ uint64_t s_item;
    const uint32_t l = threadIdx.x, rc = l & 31u, kh = l >> 5;
    constexpr double kUnscale = 1.0 / ((double)kFacScale * (double)kFacScale);
    for (uint64_t item = first_item(a, &s_item); item < a.total_items;
         item = next_item(a, &s_item, item)) {
        // (object, chunk, slice); slices hold a multiple of 16 observations
        uint64_t oc = item;
        uint32_t slice = 0, obj, chunk;
        if (a.slices > 1) {
            oc = item / a.slices;
            slice = (uint32_t)(item - oc * a.slices);
        }
        item_coords(a, oc, obj, chunk);
        const uint64_t obeg = a.offsets[obj], oend = a.offsets[obj + 1];
        uint64_t beg = obeg, end = oend;
        if (a.slices > 1) {
            const uint64_t len = (((oend - obeg) + a.slices - 1) / a.slices + 15u) & ~(uint64_t)15u;
            const uint64_t b = obeg + (uint64_t)slice * len;
            beg = b < oend ? b : oend;
            end = (beg + len) < oend ? beg + len : oend;
        }
        const float ys = a.obj_ys[obj];
        const uint32_t kbase = chunk * (uint32_t)(T * M);
        const double fa = a.fmin_rev + (double)(kbase + 32u * rc) * a.df_rev;  // row rc, column 0
        const uint4 *cf = a.colfac +
                          (colfac_slot(obeg, obj) + ((beg - obeg) >> 4)) * (uint64_t)(kColfacFrags * 64) + l;
        f32x16_t z1r = {}, z1r_x = {}, z1i = {}, z1i_x = {};  // Z1: head x head, cross products
        f32x16_t z2r = {}, z2i = {};
        // after every kFoldObs observations the accumulators move into FP32
        // sums in LDS (each lane its own entries, so no barrier)
        bool folded = false;
        auto fold = [&]() {
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                s_acc[0][v][l] = (folded ? s_acc[0][v][l] : 0.0f) + (z1r[v] + z1r_x[v]);
                s_acc[1][v][l] = (folded ? s_acc[1][v][l] : 0.0f) + (z1i[v] + z1i_x[v]);
                s_acc[2][v][l] = (folded ? s_acc[2][v][l] : 0.0f) + z2r[v];
                s_acc[3][v][l] = (folded ? s_acc[3][v][l] : 0.0f) + z2i[v];
            }
            z1r = f32x16_t{};
            z1r_x = f32x16_t{};
            z1i = f32x16_t{};
            z1i_x = f32x16_t{};
            z2r = f32x16_t{};
            z2i = f32x16_t{};
            folded = true;
        };
        for (uint64_t g0 = beg; g0 < end; g0 += kFoldObs) {
        const uint64_t gend = (end - g0) < kFoldObs ? end : g0 + kFoldObs;
        for (uint64_t b0 = g0; b0 < gend; b0 += T) {
            const uint32_t cnt = (uint32_t)((gend - b0) < (uint64_t)T ? (gend - b0) : (uint64_t)T);
            __syncthreads();  // also orders the previous item's peak merge before this one's
            if (l < cnt) {
                s_t[l] = a.rec_t[b0 + l];
                s_y[l] = a.rec_a[b0 + l].x * (ys * kFacScale);
            }
            __syncthreads();
            for (uint32_t j0 = 0; j0 < cnt; j0 += 16, cf += kColfacFrags * 64) {
                half8_t col[kColfacFrags];
#pragma unroll
                for (int i = 0; i < kColfacFrags; ++i) {
                    const uint4 u = cf[i * 64];
                    __builtin_memcpy(&col[i], &u, sizeof(u));
                }
                half8_t ar_h, ar_l, ai_h, ai_l, qr_h, qr_l, qi_h, qi_l;  // y u and u^2
#pragma unroll
                for (int v = 0; v < 8; ++v) {
                    const uint32_t j = j0 + 8u * kh + (uint32_t)v;  // < T: j0 <= T - 16
                    const bool ok = j < cnt;
                    const double tj = s_t[j];
                    const float yj = ok ? s_y[j] : 0.0f;
                    double ph = fa * tj;
                    ph -= __builtin_rint(ph);
                    float c = __builtin_amdgcn_cosf((float)ph);
                    float s = __builtin_amdgcn_sinf((float)ph);
                    c = ok ? c : 0.0f;
                    s = ok ? s : 0.0f;
                    _Float16 h, t;
                    split16(yj * c, h, t);
                    ar_h[v] = h;
                    ar_l[v] = t;
                    split16(yj * s, h, t);
                    ai_h[v] = h;
                    ai_l[v] = t;
                    const float cc = c * c, ss = s * s, g = kFacScale * (2.0f - (cc + ss));  // u^2/|u|^2
                    split16((cc - ss) * g, h, t);
                    qr_h[v] = h;
                    qr_l[v] = t;
                    split16(2.0f * c * s * g, h, t);
                    qi_h[v] = h;
                    qi_l[v] = t;
                }
                // Re Z1 = (y u)_r v_r - (y u)_i v_i,  Im Z1 = (y u)_r v_i + (y u)_i v_r
                LS_MFMA(z1r, ar_h, col[0]);
                LS_MFMA(z1r_x, ar_h, col[1]);
                LS_MFMA(z1r_x, ar_l, col[0]);
                LS_MFMA(z1r, ai_h, col[4]);
                LS_MFMA(z1r_x, ai_h, col[5]);
                LS_MFMA(z1r_x, ai_l, col[4]);
                LS_MFMA(z1i, ar_h, col[2]);
                LS_MFMA(z1i_x, ar_h, col[3]);
                LS_MFMA(z1i_x, ar_l, col[2]);
                LS_MFMA(z1i, ai_h, col[0]);
                LS_MFMA(z1i_x, ai_h, col[1]);
                LS_MFMA(z1i_x, ai_l, col[0]);
                // the same for Z2 with u^2 and w
                LS_MFMA(z2r, qr_h, col[6]);
                LS_MFMA(z2r, qr_h, col[7]);
                LS_MFMA(z2r, qr_l, col[6]);
                LS_MFMA(z2r, qi_h, col[10]);
                LS_MFMA(z2r, qi_h, col[11]);
                LS_MFMA(z2r, qi_l, col[10]);
                LS_MFMA(z2i, qr_h, col[8]);
                LS_MFMA(z2i, qr_h, col[9]);
                LS_MFMA(z2i, qr_l, col[8]);
                LS_MFMA(z2i, qi_h, col[6]);
                LS_MFMA(z2i, qi_h, col[7]);
                LS_MFMA(z2i, qi_l, col[6]);
            }
        }
        if (gend < end) fold();
        }
        // the item's FP64 sums of result entry v
        auto sums = [&](int v, double &XC, double &XS, double &C2, double &S2) {
            XC = (double)z1r[v] + (double)z1r_x[v];
            XS = (double)z1i[v] + (double)z1i_x[v];
            C2 = (double)z2r[v];
            S2 = (double)z2i[v];
            if (folded) {
                XC += (double)s_acc[0][v][l];
                XS += (double)s_acc[1][v][l];
                C2 += (double)s_acc[2][v][l];
                S2 += (double)s_acc[3][v][l];
            }
            XC *= kUnscale;
            XS *= kUnscale;
            C2 *= kUnscale;
            S2 *= kUnscale;
        };
        const double iys = 1.0 / (double)ys;  // a power of two
        if (a.slices > 1) {
            // the slice's FP64 sums in the layout k_combine_slices reads:
            // XC, XS, CC, CS of bin m T + tid at [(q M + m) T + tid]
            double *dst = a.part_sums + (((uint64_t)obj * a.chunks + chunk) * a.slices + slice) *
                                            (uint64_t)(4 * M * T);
            const double ns = (double)(end - beg);
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const uint32_t row = 8u * (uint32_t)(v / 4) + 4u * kh + (uint32_t)(v % 4);
                const uint32_t bl = 32u * row + rc, m = bl / T, tt = bl % T;
                double XC, XS, C2, S2;
                sums(v, XC, XS, C2, S2);
                dst[(0 * M + m) * T + tt] = XC * iys;
                dst[(1 * M + m) * T + tt] = XS * iys;
                dst[(2 * M + m) * T + tt] = 0.5 * (ns + C2);
                dst[(3 * M + m) * T + tt] = 0.5 * S2;
            }
            continue;
        }
        const double scale = a.obj_scale[obj] * iys * iys;
        const double nobs = (double)(oend - obeg);
        OUT *pg = static_cast<OUT *>(a.pgram);
        Peaks L;
        peaks_init(L);
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const uint32_t row = 8u * (uint32_t)(v / 4) + 4u * kh + (uint32_t)(v % 4);
            const uint32_t kf = kbase + 32u * row + rc;  // ascending in v
            double XC, XS, C2, S2;
            sums(v, XC, XS, C2, S2);
            const double CC = 0.5 * (nobs + C2), SS = nobs - CC, CS = 0.5 * S2;
            const double D = CC * SS - CS * CS;
            const double P = 0.5 * (SS * XC * XC - 2.0 * CS * XC * XS + CC * XS * XS) / D;
            const OUT pw = (OUT)(P * scale);
            if (kf < a.nf) {
                if (pg) pg[(uint64_t)obj * a.nf + kf] = pw;
                peaks_add(L, a.top_k, (double)pw, kf);
            }
        }
        if (a.top_k) write_partial<OUT, T>(a, L, obj, chunk, s_pp, s_pi);
    }
}


// Generalized variant of the matrix-core tile.  With q = sqrt(w / W) (the
// records' normalised weights) and y_c = y - ybar, three complex sums per bin:
//   Zw = sum q^2 e^{i phi},  Zy = sum q (q y_c) e^{i phi},  Z2 = sum q^2 e^{2 i phi}
// give the seven weighted sums of the generalized power: C, S = Zw;
// YC, YS = Zy; CC, SS = (Wf +- Re Z2) / 2; CS = Im Z2 / 2, with Wf = sum q^2.
// 36 MFMAs and 12 accumulators (head x head, cross products) per 16
// observations.  The fold keeps Zy in FP64 and the weight-only Zw and Z2 in
// FP32 (34 KiB of LDS: four workgroups per CU, one per SIMD).
template <typename OUT>
__global__ __launch_bounds__(64) void k_sweep_mfma_gen(SweepArgs a) {
    constexpr int T = 64, M = 16;
    constexpr uint32_t kFoldObs = 256;
    __shared__ double s_t[T];
    __shared__ float s_q[T], s_y[T];
    __shared__ double s_accy[2][16][T];  // FP64 Re/Im of Zy of a lane's 16 bins
    __shared__ float s_accw[4][16][T];   // FP32 Re/Im of Zw and Z2 (weights only)
    __shared__ double s_pp[1][kPeaks];
    __shared__ uint32_t s_pi[1][kPeaks];
    __shared__ uint64_t s_item;
    const uint32_t l = threadIdx.x, rc = l & 31u, kh = l >> 5;
    constexpr double kUnscale = 1.0 / ((double)kFacScale * (double)kFacScale);
    for (uint64_t item = first_item(a, &s_item); item < a.total_items;
         item = next_item(a, &s_item, item)) {
        uint64_t oc = item;
        uint32_t slice = 0, obj, chunk;
        if (a.slices > 1) {
            oc = item / a.slices;
            slice = (uint32_t)(item - oc * a.slices);
        }
        item_coords(a, oc, obj, chunk);
        const uint64_t obeg = a.offsets[obj], oend = a.offsets[obj + 1];
        uint64_t beg = obeg, end = oend;
        if (a.slices > 1) {
            const uint64_t len = (((oend - obeg) + a.slices - 1) / a.slices + 15u) & ~(uint64_t)15u;
            const uint64_t b = obeg + (uint64_t)slice * len;
            beg = b < oend ? b : oend;
            end = (beg + len) < oend ? beg + len : oend;
        }
        const float ys = a.obj_ys[obj];
        const uint32_t kbase = chunk * (uint32_t)(T * M);
        const double fa = a.fmin_rev + (double)(kbase + 32u * rc) * a.df_rev;
        const uint4 *cf = a.colfac +
                          (colfac_slot(obeg, obj) + ((beg - obeg) >> 4)) * (uint64_t)(kColfacFrags * 64) + l;
        f32x16_t zw[4] = {}, zy[4] = {}, z2[4] = {};  // {Re hh, Re x, Im hh, Im x}
        double wq = 0.0;                             // this lane's share of sum q^2 (slices)
        bool folded = false;
        auto fold = [&]() {
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                float pw[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) pw[q] = folded ? s_accw[q][v][l] : 0.0f;
                const double py0 = folded ? s_accy[0][v][l] : 0.0, py1 = folded ? s_accy[1][v][l] : 0.0;
                s_accw[0][v][l] = pw[0] + (zw[0][v] + zw[1][v]);
                s_accw[1][v][l] = pw[1] + (zw[2][v] + zw[3][v]);
                s_accy[0][v][l] = py0 + (double)zy[0][v] + (double)zy[1][v];
                s_accy[1][v][l] = py1 + (double)zy[2][v] + (double)zy[3][v];
                s_accw[2][v][l] = pw[2] + (z2[0][v] + z2[1][v]);
                s_accw[3][v][l] = pw[3] + (z2[2][v] + z2[3][v]);
            }
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                zw[i] = f32x16_t{};
                zy[i] = f32x16_t{};
                z2[i] = f32x16_t{};
            }
            folded = true;
        };
        for (uint64_t g0 = beg; g0 < end; g0 += kFoldObs) {
        const uint64_t gend = (end - g0) < kFoldObs ? end : g0 + kFoldObs;
        for (uint64_t b0 = g0; b0 < gend; b0 += T) {
            const uint32_t cnt = (uint32_t)((gend - b0) < (uint64_t)T ? (gend - b0) : (uint64_t)T);
            __syncthreads();  // also orders the previous item's peak merge before this one's
            if (l < cnt) {
                const float4 r = a.rec_a[b0 + l];
                s_t[l] = a.rec_t[b0 + l];
                s_q[l] = r.y;
                s_y[l] = r.x * (ys * kFacScale);
                wq += (double)(r.y * r.y);
            }
            __syncthreads();
            for (uint32_t j0 = 0; j0 < cnt; j0 += 16, cf += kColfacFrags * 64) {
                half8_t col[kColfacFrags];
#pragma unroll
                for (int i = 0; i < kColfacFrags; ++i) {
                    const uint4 u = cf[i * 64];
                    __builtin_memcpy(&col[i], &u, sizeof(u));
                }
                half8_t wr_h, wr_l, wi_h, wi_l, yr_h, yr_l, yi_h, yi_l, qr_h, qr_l, qi_h, qi_l;
#pragma unroll
                for (int v = 0; v < 8; ++v) {
                    const uint32_t j = j0 + 8u * kh + (uint32_t)v;  // < T: j0 <= T - 16
                    const bool ok = j < cnt;
                    const double tj = s_t[j];
                    const float qj = ok ? s_q[j] : 0.0f, yj = ok ? s_y[j] : 0.0f;
                    double ph = fa * tj;
                    ph -= __builtin_rint(ph);
                    float c = __builtin_amdgcn_cosf((float)ph);
                    float s = __builtin_amdgcn_sinf((float)ph);
                    c = ok ? c : 0.0f;
                    s = ok ? s : 0.0f;
                    const float q2 = qj * qj * kFacScale, qy = qj * yj;
                    _Float16 h, t;
                    split16(q2 * c, h, t);
                    wr_h[v] = h;
                    wr_l[v] = t;
                    split16(q2 * s, h, t);
                    wi_h[v] = h;
                    wi_l[v] = t;
                    split16(qy * c, h, t);
                    yr_h[v] = h;
                    yr_l[v] = t;
                    split16(qy * s, h, t);
                    yi_h[v] = h;
                    yi_l[v] = t;
                    const float cc = c * c, ss = s * s, g = q2 * (2.0f - (cc + ss));  // q^2 u^2/|u|^2
                    split16((cc - ss) * g, h, t);
                    qr_h[v] = h;
                    qr_l[v] = t;
                    split16(2.0f * c * s * g, h, t);
                    qi_h[v] = h;
                    qi_l[v] = t;
                }
                // Re Z = p_r f_r - p_i f_i,  Im Z = p_r f_i + p_i f_r; column
                // fragments f = v (col 0..5) for Zw, Zy and w = v^2 (col 6..11) for Z2
                LS_MFMA(zw[0], wr_h, col[0]);
                LS_MFMA(zw[1], wr_h, col[1]);
                LS_MFMA(zw[1], wr_l, col[0]);
                LS_MFMA(zw[0], wi_h, col[4]);
                LS_MFMA(zw[1], wi_h, col[5]);
                LS_MFMA(zw[1], wi_l, col[4]);
                LS_MFMA(zw[2], wr_h, col[2]);
                LS_MFMA(zw[3], wr_h, col[3]);
                LS_MFMA(zw[3], wr_l, col[2]);
                LS_MFMA(zw[2], wi_h, col[0]);
                LS_MFMA(zw[3], wi_h, col[1]);
                LS_MFMA(zw[3], wi_l, col[0]);
                LS_MFMA(zy[0], yr_h, col[0]);
                LS_MFMA(zy[1], yr_h, col[1]);
                LS_MFMA(zy[1], yr_l, col[0]);
                LS_MFMA(zy[0], yi_h, col[4]);
                LS_MFMA(zy[1], yi_h, col[5]);
                LS_MFMA(zy[1], yi_l, col[4]);
                LS_MFMA(zy[2], yr_h, col[2]);
                LS_MFMA(zy[3], yr_h, col[3]);
                LS_MFMA(zy[3], yr_l, col[2]);
                LS_MFMA(zy[2], yi_h, col[0]);
                LS_MFMA(zy[3], yi_h, col[1]);
                LS_MFMA(zy[3], yi_l, col[0]);
                LS_MFMA(z2[0], qr_h, col[6]);
                LS_MFMA(z2[1], qr_h, col[7]);
                LS_MFMA(z2[1], qr_l, col[6]);
                LS_MFMA(z2[0], qi_h, col[10]);
                LS_MFMA(z2[1], qi_h, col[11]);
                LS_MFMA(z2[1], qi_l, col[10]);
                LS_MFMA(z2[2], qr_h, col[8]);
                LS_MFMA(z2[3], qr_h, col[9]);
                LS_MFMA(z2[3], qr_l, col[8]);
                LS_MFMA(z2[2], qi_h, col[6]);
                LS_MFMA(z2[3], qi_h, col[7]);
                LS_MFMA(z2[3], qi_l, col[6]);
            }
        }
        if (gend < end) fold();
        }
        // the item's FP64 sums of result entry v: Re/Im Zw, Zy (y unscaled), Z2
        auto sums = [&](int v, double (&z)[6]) {
            z[0] = (double)zw[0][v] + (double)zw[1][v];
            z[1] = (double)zw[2][v] + (double)zw[3][v];
            z[2] = (double)zy[0][v] + (double)zy[1][v];
            z[3] = (double)zy[2][v] + (double)zy[3][v];
            z[4] = (double)z2[0][v] + (double)z2[1][v];
            z[5] = (double)z2[2][v] + (double)z2[3][v];
            if (folded) {
#pragma unroll
                for (int q = 0; q < 2; ++q) {
                    z[q] += s_accw[q][v][l];
                    z[2 + q] += s_accy[q][v][l];
                    z[4 + q] += s_accw[2 + q][v][l];
                }
            }
            const double iys = 1.0 / (double)ys;
#pragma unroll
            for (int q = 0; q < 6; ++q) z[q] *= kUnscale * (q == 2 || q == 3 ? iys : 1.0);
        };
        if (a.slices > 1) {
            // the slice's seven FP64 sums in the layout k_combine_slices reads,
            // with the slice's weight sum for CC and SS
            double wsum = wq;
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) wsum += __shfl_xor(wsum, off, 64);
            double *dst = a.part_sums + (((uint64_t)obj * a.chunks + chunk) * a.slices + slice) *
                                            (uint64_t)(7 * M * T);
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const uint32_t row = 8u * (uint32_t)(v / 4) + 4u * kh + (uint32_t)(v % 4);
                const uint32_t bl = 32u * row + rc, m = bl / T, tt = bl % T;
                double z[6];
                sums(v, z);
                dst[(0 * M + m) * T + tt] = z[0];
                dst[(1 * M + m) * T + tt] = z[1];
                dst[(2 * M + m) * T + tt] = z[2];
                dst[(3 * M + m) * T + tt] = z[3];
                dst[(4 * M + m) * T + tt] = 0.5 * (wsum + z[4]);
                dst[(5 * M + m) * T + tt] = 0.5 * (wsum - z[4]);
                dst[(6 * M + m) * T + tt] = 0.5 * z[5];
            }
            continue;
        }
        const double scale = a.obj_scale[obj];
        const double Wf = a.obj_gen[kGen * obj + 0], Yb = a.obj_gen[kGen * obj + 1];
        const double YY = a.obj_gen[kGen * obj + 2], iW = 1.0 / Wf;
        OUT *pg = static_cast<OUT *>(a.pgram);
        Peaks L;
        peaks_init(L);
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const uint32_t row = 8u * (uint32_t)(v / 4) + 4u * kh + (uint32_t)(v % 4);
            const uint32_t kf = kbase + 32u * row + rc;  // ascending in v
            double z[6];
            sums(v, z);
            const double C = z[0] * iW, S = z[1] * iW;
            const double YC = z[2] * iW - Yb * C, YS = z[3] * iW - Yb * S;
            const double CC = 0.5 * (Wf + z[4]) * iW - C * C;
            const double SS = 0.5 * (Wf - z[4]) * iW - S * S;
            const double CS = 0.5 * z[5] * iW - C * S;
            const double D = CC * SS - CS * CS;
            const double P = (SS * YC * YC - 2.0 * CS * YC * YS + CC * YS * YS) / (D * YY);
            const OUT pw = (OUT)(P * scale);
            if (kf < a.nf) {
                if (pg) pg[(uint64_t)obj * a.nf + kf] = pw;
                peaks_add(L, a.top_k, (double)pw, kf);
            }
        }
        if (a.top_k) write_partial<OUT, T>(a, L, obj, chunk, s_pp, s_pi);
    }
}

}  // namespace

#define LS_SPLIT32(GEN, OUTT, DT, TT, MM)                                      \
    {LSAMD_FORM_SPLIT32, GEN, DT, TT, MM, k_sweep_split32<(GEN) != 0, OUTT, TT, MM>, \
     k_combine_slices<true, (GEN) != 0, OUTT, TT, MM>}
#define LS_SPLIT(GEN, OUTT, DT, TT, MM)                                      \
    {LSAMD_FORM_SPLIT, GEN, DT, TT, MM, k_sweep_split64<(GEN) != 0, OUTT, TT, MM>, \
     k_combine_slices<false, (GEN) != 0, OUTT, TT, MM>}

const TileEntry kSplitTiles[] = {
    // split32 on the matrix cores (standard variant): 64 threads, 32 x 32 bins
    {LSAMD_FORM_SPLIT32, 0, LSAMD_F32, 64, 16, k_sweep_mfma<float>,
     k_combine_slices<true, false, float, 64, 16>, k_colfac},
    {LSAMD_FORM_SPLIT32, 0, LSAMD_F64, 64, 16, k_sweep_mfma<double>,
     k_combine_slices<true, false, double, 64, 16>, k_colfac},
    {LSAMD_FORM_SPLIT32, 1, LSAMD_F32, 64, 16, k_sweep_mfma_gen<float>,
     k_combine_slices<true, true, float, 64, 16>, k_colfac},
    {LSAMD_FORM_SPLIT32, 1, LSAMD_F64, 64, 16, k_sweep_mfma_gen<double>,
     k_combine_slices<true, true, double, 64, 16>, k_colfac},
    // split form (FP64 rotation and sums), both output types
    LS_SPLIT(0, float, LSAMD_F32, 256, 8), LS_SPLIT(0, double, LSAMD_F64, 256, 8),
    LS_SPLIT(0, float, LSAMD_F32, 256, 4), LS_SPLIT(0, double, LSAMD_F64, 256, 4),
    LS_SPLIT(0, float, LSAMD_F32, 256, 16), LS_SPLIT(0, double, LSAMD_F64, 256, 16),
    LS_SPLIT(0, float, LSAMD_F32, 128, 8), LS_SPLIT(0, double, LSAMD_F64, 128, 8),
    LS_SPLIT(0, float, LSAMD_F32, 512, 8), LS_SPLIT(0, double, LSAMD_F64, 512, 8),
    LS_SPLIT(1, float, LSAMD_F32, 256, 4), LS_SPLIT(1, double, LSAMD_F64, 256, 4),
    LS_SPLIT(1, float, LSAMD_F32, 256, 8), LS_SPLIT(1, double, LSAMD_F64, 256, 8),
    LS_SPLIT(1, float, LSAMD_F32, 128, 8), LS_SPLIT(1, double, LSAMD_F64, 128, 8),
    LS_SPLIT(0, float, LSAMD_F32, 128, 16), LS_SPLIT(0, double, LSAMD_F64, 128, 16),
    LS_SPLIT(1, float, LSAMD_F32, 256, 2), LS_SPLIT(1, double, LSAMD_F64, 256, 2),
    // split32 form (FP32 rotation and chunk sums)
    LS_SPLIT32(0, float, LSAMD_F32, 256, 8), LS_SPLIT32(0, double, LSAMD_F64, 256, 8),
    LS_SPLIT32(0, float, LSAMD_F32, 256, 4), LS_SPLIT32(0, double, LSAMD_F64, 256, 4),
    LS_SPLIT32(0, float, LSAMD_F32, 256, 2), LS_SPLIT32(0, double, LSAMD_F64, 256, 2),
    LS_SPLIT32(0, float, LSAMD_F32, 128, 4), LS_SPLIT32(0, double, LSAMD_F64, 128, 4),
    LS_SPLIT32(0, float, LSAMD_F32, 512, 4), LS_SPLIT32(0, double, LSAMD_F64, 512, 4),
    LS_SPLIT32(1, float, LSAMD_F32, 256, 4), LS_SPLIT32(1, double, LSAMD_F64, 256, 4),
    LS_SPLIT32(1, float, LSAMD_F32, 256, 2), LS_SPLIT32(1, double, LSAMD_F64, 256, 2),
    LS_SPLIT32(1, float, LSAMD_F32, 128, 4), LS_SPLIT32(1, double, LSAMD_F64, 128, 4),
    LS_SPLIT32(0, float, LSAMD_F32, 256, 6), LS_SPLIT32(0, double, LSAMD_F64, 256, 6),
    LS_SPLIT32(1, float, LSAMD_F32, 256, 3), LS_SPLIT32(1, double, LSAMD_F64, 256, 3),
    LS_SPLIT(0, float, LSAMD_F32, 256, 12), LS_SPLIT(0, double, LSAMD_F64, 256, 12),
    LS_SPLIT(0, float, LSAMD_F32, 256, 14), LS_SPLIT(0, double, LSAMD_F64, 256, 14),
    LS_SPLIT32(0, float, LSAMD_F32, 128, 6), LS_SPLIT32(0, double, LSAMD_F64, 128, 6),
    LS_SPLIT32(0, float, LSAMD_F32, 128, 8), LS_SPLIT32(0, double, LSAMD_F64, 128, 8),
    LS_SPLIT32(0, float, LSAMD_F32, 64, 6), LS_SPLIT32(0, double, LSAMD_F64, 64, 6),
    LS_SPLIT32(1, float, LSAMD_F32, 256, 6), LS_SPLIT32(1, double, LSAMD_F64, 256, 6),
    LS_SPLIT32(1, float, LSAMD_F32, 128, 3), LS_SPLIT32(1, double, LSAMD_F64, 128, 3),
};
const int kSplitTileCount = (int)(sizeof(kSplitTiles) / sizeof(kSplitTiles[0]));

}  // namespace lsamd
