// Split-form sweep kernels (split: FP64 rotation and sums; split32: FP32
// rotation and FP32 partial sums folded into FP64, and the matrix-core
// tiles), the observation-slice combiner, and their tile registry entries.

#include "ls_device.h"

namespace lsamd {

namespace {

// ---------------------------------------------------------------------------
// FP64 recomputation of ill-conditioned bins.  Every split-form power comes
// from a 2x2 (standard) or centred 2x2 (generalized) least-squares system; its
// relative error is the sums' relative error times a condition number that
// the sums themselves give:
//   standard:    k = (CC + SS)^2 / (4 D)                 (1 for balanced phases)
//   generalized: k = (1 + C^2 + S^2) (CC + SS) / D      (centred CC, SS, CS;
//                the 1 + C^2 + S^2 covers the centring CC = CCw - C^2)
// with D = CC SS - CS^2.  The FP32 sums (split32) leave ~2.4e-7 k relative
// (measured on the 1000-object generalized batch at the 1 cycle/day alias,
// k ~ 110-230), the FP32-seeded FP64 sums (split) less.  Bins with
// k > kCondMax (a few per 10^6 for nightly-sampled objects) and every bin of
// objects of at most kSmallObs observations (5 points against the
// generalized 3-parameter model can give k ~ 10^5) are computed again from
// the caller's FP64 t, y, dy with FP64 phasors.
// ---------------------------------------------------------------------------
constexpr double kCondMax = 16.0;
constexpr uint32_t kSmallObs = 16;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Lane src's v, in every lane (src the same in all lanes).
__device__ __forceinline__ double lane_value(double v, int src) {
    const unsigned long long u = (unsigned long long)__double_as_longlong(v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, src);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), src);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// sin and cos of 2 pi r for |r| <= 1/2 revolution: the octant q = rint(4 r)
// is taken off exactly, and x = 2 pi (r - q/4), |x| <= pi/4, goes through the
// Taylor series of sin to x^17 and of cos to x^16 (the first omitted terms
// are below 5e-17 and 2e-18), in Horner form.
__device__ __forceinline__ void sincos_rev64(double r, double &s, double &c) {
    const double q = __builtin_rint(4.0 * r);
    const double x = (r - 0.25 * q) * kTwoPi, x2 = x * x;
    double ps = -1.0 / 355687428096000.0;  // -1/17!
    ps = fma(ps, x2, 1.0 / 1307674368000.0);
    ps = fma(ps, x2, -1.0 / 6227020800.0);
    ps = fma(ps, x2, 1.0 / 39916800.0);
    ps = fma(ps, x2, -1.0 / 362880.0);
    ps = fma(ps, x2, 1.0 / 5040.0);
    ps = fma(ps, x2, -1.0 / 120.0);
    ps = fma(ps, x2, 1.0 / 6.0);
    const double sx = fma(-ps * x2, x, x);  // x - x^3/3! + ... - x^17/17!... with alternating signs
    double pc = 1.0 / 20922789888000.0;     // 1/16!
    pc = fma(pc, x2, -1.0 / 87178291200.0);
    pc = fma(pc, x2, 1.0 / 479001600.0);
    pc = fma(pc, x2, -1.0 / 3628800.0);
    pc = fma(pc, x2, 1.0 / 40320.0);
    pc = fma(pc, x2, -1.0 / 720.0);
    pc = fma(pc, x2, 1.0 / 24.0);
    pc = fma(pc, x2, -0.5);
    const double cx = fma(pc, x2, 1.0);
    const int o = (int)q & 3;  // sin/cos of q pi/2 + x
    s = o == 0 ? sx : o == 1 ? cx : o == 2 ? -sx : -cx;
    c = o == 0 ? cx : o == 1 ? -sx : o == 2 ? -cx : sx;
}

// e^{2 pi i f t} in FP64: the phase reduced to [-1/2, 1/2] revolutions with
// the product f t fused (exact before the subtraction).
__device__ __forceinline__ void phasor64(double f, double t, double &c, double &s) {
    const double n = __builtin_rint(f * t);
    sincos_rev64(fma(f, t, -n), s, c);
}

// n / d for d > 0 (normal): the hardware reciprocal refined by two Newton
// steps, then one correction of the quotient (within an ulp or two; the
// IEEE division sequence is twice as long).
__device__ __forceinline__ double fdiv(double n, double d) {
    double r = __builtin_amdgcn_rcp(d);
    r = fma(fma(-d, r, 1.0), r, r);
    r = fma(fma(-d, r, 1.0), r, r);
    const double q = n * r;
    return fma(fma(-d, q, n), r, q);
}

// Power of the standard variant from its uncentred sums (SS = N - CC);
// `ill` when its condition number exceeds kCondMax (or D <= 0, or NaN).
__device__ __forceinline__ double std_power(double XC, double XS, double CC, double CS, double N, bool &ill) {
    const double SS = N - CC, D = CC * SS - CS * CS;
    ill = !((CC + SS) * (CC + SS) <= 4.0 * kCondMax * D) || !(D > 0.0);
    return fdiv(0.5 * (SS * XC * XC - 2.0 * CS * XC * XS + CC * XS * XS), D);
}

// Power of the generalized variant from sums over weights normalised to
// sum 1 (C, S, CCw, SSw, CSw, YCw, YSw; Yb = weighted mean of the y in them,
// YY the centred weighted sum of squares); `ill` as for std_power.
__device__ __forceinline__ double gen_power(double C, double S, double CCw, double SSw, double CSw,
                                            double YCw, double YSw, double Yb, double YY, bool &ill) {
    const double YC = YCw - Yb * C, YS = YSw - Yb * S;
    const double CC = CCw - C * C, SS = SSw - S * S, CS = CSw - C * S;
    const double D = CC * SS - CS * CS;
    ill = !((1.0 + C * C + S * S) * (CC + SS) <= kCondMax * D) || !(D > 0.0);
    return fdiv(SS * YC * YC - 2.0 * CS * YC * YS + CC * YS * YS, D * YY);
}

// FP64 power of bin kf of object obj (before the object's scale); every lane
// of the wavefront calls it with the same (obj, kf) and gets the result.
// Standard: raw y, unit weights.  Generalized: w = dy^-2 / W and y - ybar with
// W and ybar from the statistics kernel (the power does not depend on them
// beyond rounding: the weights are normalised and the mean floats).
template <bool GEN, typename IN>
__device__ double exact_power(const SweepArgs &a, uint32_t obj, uint32_t kf) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t b = a.offsets[obj], e = a.offsets[obj + 1];
    const double f = a.fmin_rev + (double)kf * a.df_rev;
    const IN *y = static_cast<const IN *>(a.y);
    const IN *dy = static_cast<const IN *>(a.dy);
    double iW = 1.0, ybar = 0.0;
    if constexpr (GEN) {
        iW = 1.0 / a.obj_gen[kGen * obj + 4];
        ybar = a.obj_gen[kGen * obj + 3];
    }
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0, s4 = 0.0, s5 = 0.0, s6 = 0.0, s7 = 0.0, s8 = 0.0;
    for (uint64_t j = b + lane; j < e; j += 64) {
        double c, sn;
        phasor64(f, a.rec_t[j], c, sn);
        if constexpr (GEN) {
            const double dj = (double)dy[j], w = iW / (dj * dj), yc = (double)y[j] - ybar;
            const double wc = w * c, wy = w * yc;
            s0 += wc;
            s1 = fma(w, sn, s1);
            s2 = fma(wc, c, s2);
            s3 = fma(wc, sn, s3);
            s4 = fma(wy, c, s4);
            s5 = fma(wy, sn, s5);
            s6 += w;
            s7 += wy;
            s8 = fma(wy, yc, s8);
        } else {
            const double yj = (double)y[j];
            s0 = fma(yj, c, s0);
            s1 = fma(yj, sn, s1);
            s2 = fma(c, c, s2);
            s3 = fma(c, sn, s3);
        }
    }
    bool ill;
    if constexpr (GEN) {
        const double W = wave_sum(s6), iWn = 1.0 / W;
        const double C = wave_sum(s0) * iWn, S = wave_sum(s1) * iWn, CCw = wave_sum(s2) * iWn;
        const double CSw = wave_sum(s3) * iWn, YCw = wave_sum(s4) * iWn, YSw = wave_sum(s5) * iWn;
        const double Yb = wave_sum(s7) * iWn, YY = wave_sum(s8) * iWn - Yb * Yb;
        return gen_power(C, S, CCw, 1.0 - CCw, CSw, YCw, YSw, Yb, YY, ill);
    } else {
        return std_power(wave_sum(s0), wave_sum(s1), wave_sum(s2), wave_sum(s3), (double)(e - b), ill);
    }
}

// Recomputes, in FP64, the bins this wavefront's lanes flagged (bit m of a
// lane's `flags` for its bin m, whose number bin_of(lane, m) gives), one bin
// at a time with the whole wavefront.  One loop over the flagged (lane, bin)
// pairs, so that exact_power is inlined once.
template <bool GEN, typename IN, int M, typename BinOf>
__device__ __forceinline__ void recompute_flagged(const SweepArgs &a, uint32_t obj, uint32_t flags, BinOf bin_of,
                                                  double (&P)[M]) {
    const uint32_t lane = threadIdx.x & 63u;
    uint64_t lanes = __ballot(flags != 0u);
    while (lanes) {
        const uint32_t src = (uint32_t)__builtin_ctzll(lanes);
        const uint32_t f = (uint32_t)__shfl((int)flags, (int)src, 64);
        const int m = __builtin_ctz(f);
        const double pe = exact_power<GEN, IN>(a, obj, bin_of(src, m));
        if (lane == src) {
#pragma unroll
            for (int q = 0; q < M; ++q) P[q] = q == m ? pe : P[q];
            flags &= flags - 1u;
        }
        lanes = __ballot(flags != 0u);
    }
}

// FP64 powers (before the object's scale) of a lane's 16 bins in the
// matrix-core tile's layout, bin kbase + rc + 32 (8 (v / 4) + 4 kh + v % 4),
// for an object of at most kSmallObs observations: per observation one FP64
// phasor at the lane's first bin and FP64 rotations by 32 and 256 bins
// (sums as exact_power's).  All 64 lanes call it (the rotations are read
// across lanes).
template <bool GEN, typename IN>
__device__ void small_object_powers(const SweepArgs &a, uint32_t obj, uint32_t kbase, uint32_t rc,
                                    uint32_t kh, double (&P)[16]) {
    constexpr int NS = GEN ? 6 : 4;  // generalized C, S, CCw, CSw, YCw, YSw; standard XC, XS, CC, CS
    const uint64_t b = a.offsets[obj], e = a.offsets[obj + 1];
    const double f0 = a.fmin_rev + (double)(kbase + rc + 128u * kh) * a.df_rev;
    const double f32 = 32.0 * a.df_rev, f256 = 256.0 * a.df_rev;
    const IN *y = static_cast<const IN *>(a.y);
    const IN *dy = static_cast<const IN *>(a.dy);
    double iW = 1.0, ybar = 0.0;
    if constexpr (GEN) {
        iW = 1.0 / a.obj_gen[kGen * obj + 4];
        ybar = a.obj_gen[kGen * obj + 3];
    }
    double acc[NS][16];
#pragma unroll
    for (int q = 0; q < NS; ++q)
#pragma unroll
        for (int v = 0; v < 16; ++v) acc[q][v] = 0.0;
    // The rotations by 32 and 256 bins depend on the observation alone: lane
    // i < 16 evaluates observation b + i's by 32 bins and lane 16 + i its by
    // 256, once, and the loop reads them from those lanes.
    static_assert(kSmallObs <= 16, "one lane per observation and rotation");
    const uint32_t lane = threadIdx.x & 63u, li = lane & 15u;
    double rotc = 1.0, rots = 0.0;
    if (lane < 32u && b + li < e) phasor64(lane < 16u ? f32 : f256, a.rec_t[b + li], rotc, rots);
    double W = 0.0, WY = 0.0, WYY = 0.0;
    for (uint64_t j = b; j < e; ++j) {
        const double t = a.rec_t[j];
        double w = 1.0, yv;
        if constexpr (GEN) {
            const double dj = (double)dy[j];
            w = iW / (dj * dj);
            yv = (double)y[j] - ybar;
            W += w;
            WY = fma(w, yv, WY);
            WYY = fma(w * yv, yv, WYY);
        } else {
            yv = (double)y[j];
        }
        double zc, zs;
        phasor64(f0, t, zc, zs);
        const int k = (int)(j - b);  // < kSmallObs
        const double r1c = lane_value(rotc, k), r1s = lane_value(rots, k);
        const double r8c = lane_value(rotc, 16 + k), r8s = lane_value(rots, 16 + k);
        const double wy = w * yv;
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            double c = zc, sn = zs;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int v = 4 * g + i;
                if constexpr (GEN) {
                    const double wc = w * c;
                    acc[0][v] += wc;
                    acc[1][v] = fma(w, sn, acc[1][v]);
                    acc[2][v] = fma(wc, c, acc[2][v]);
                    acc[3][v] = fma(wc, sn, acc[3][v]);
                    acc[4][v] = fma(wy, c, acc[4][v]);
                    acc[5][v] = fma(wy, sn, acc[5][v]);
                } else {
                    acc[0][v] = fma(yv, c, acc[0][v]);
                    acc[1][v] = fma(yv, sn, acc[1][v]);
                    acc[2][v] = fma(c, c, acc[2][v]);
                    acc[3][v] = fma(c, sn, acc[3][v]);
                }
                if (i < 3) {
                    const double cn = fma(c, r1c, -sn * r1s);
                    sn = fma(sn, r1c, c * r1s);
                    c = cn;
                }
            }
            if (g < 3) {
                const double cn = fma(zc, r8c, -zs * r8s);
                zs = fma(zs, r8c, zc * r8s);
                zc = cn;
            }
        }
    }
    const double N = (double)(e - b), iWn = 1.0 / W, Yb = WY * iWn, YY = WYY * iWn - Yb * Yb;
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        bool ill;
        if constexpr (GEN) {
            const double CCw = acc[2][v] * iWn;
            P[v] = gen_power(acc[0][v] * iWn, acc[1][v] * iWn, CCw, 1.0 - CCw, acc[3][v] * iWn, acc[4][v] * iWn,
                             acc[5][v] * iWn, Yb, YY, ill);
        } else {
            P[v] = std_power(acc[0][v], acc[1][v], acc[2][v], acc[3][v], N, ill);
        }
    }
}

// Bin of result entry v of lane (rc, kh) in a matrix-core work item.
__device__ __forceinline__ uint32_t mfma_bin(uint32_t kbase, uint32_t rc, uint32_t kh, int v) {
    return kbase + 32u * (8u * (uint32_t)(v / 4) + 4u * kh + (uint32_t)(v % 4)) + rc;
}

// ---------------------------------------------------------------------------
// Split forms: the powers of a thread's M bins from the object's FP64 sums
// (tau-free power, normalisation, periodogram store, the chunk's peaks).
//   standard:    XC, XS, CC, CS; SS = N - CC
//   generalized: C, S, YC, YS, CC, [SS,] CS over the normalised weights; SS
//                summed (SS_SUM, split32) or 1 - CC (split, whose unit-modulus
//                seeds make the weights sum to 1 to FP64 rounding)
// Bin m of thread tid: kbase + m T + tid, or with MFMA (sums of the matrix-core
// tile, 64 threads x 16) its result layout, mfma_bin(kbase, tid % 32, tid / 32, m).
// ---------------------------------------------------------------------------
template <bool SS_SUM, bool GEN, typename OUT, int T, int M, int NS, bool MFMA = false>
__device__ __forceinline__ void split_epilogue(const SweepArgs &a, const double (&acc)[NS][M],
                                               uint32_t obj, uint32_t chunk,
                                               double (*s_pp)[kPeaks], uint32_t (*s_pi)[kPeaks]) {
    static_assert(!MFMA || (T == 64 && M == 16), "matrix-core layout: 64 x 16");
    const uint32_t tid = threadIdx.x;
    const uint32_t kbase = chunk * (uint32_t)(T * M);
    const double scale = a.obj_scale[obj];
    const double nobs = (double)(a.offsets[obj + 1] - a.offsets[obj]);
    double Wf = 1.0, Yb = 0.0, YY = 1.0;
    if constexpr (GEN) {  // the matrix-core tile's column factors carry the FP64 weights
        Wf = a.obj_gen[kGen * obj + (MFMA ? 5 : 0)];
        Yb = a.obj_gen[kGen * obj + (MFMA ? 6 : 1)];
        YY = a.obj_gen[kGen * obj + (MFMA ? 7 : 2)];
    }
    const double iW = 1.0 / Wf;  // one division per work item, not seven per bin
    const bool small = nobs <= (double)kSmallObs;
    double P[M];
    uint32_t kf[M], flags = 0u;
#pragma unroll
    for (int m = 0; m < M; ++m) {
        kf[m] = MFMA ? mfma_bin(kbase, tid & 31u, tid >> 5, m) : kbase + m * T + tid;
        bool ill;
        if constexpr (GEN) {
            const double CCw = acc[4][m] * iW;
            const double SSw = SS_SUM ? acc[5][m] * iW : 1.0 - CCw;
            const double CSw = acc[SS_SUM ? 6 : 5][m] * iW;
            P[m] = gen_power(acc[0][m] * iW, acc[1][m] * iW, CCw, SSw, CSw, acc[2][m] * iW, acc[3][m] * iW, Yb,
                             YY, ill);
        } else {
            P[m] = std_power(acc[0][m], acc[1][m], acc[2][m], acc[3][m], nobs, ill);
        }
        if ((small || ill) && kf[m] < a.nf) flags |= 1u << m;
    }
    const uint32_t wbase = kbase + (tid & ~63u);  // bin of (lane 0, m = 0) of this wavefront
    recompute_flagged<GEN, OUT, M>(a, obj, flags, [&](uint32_t src, int m) {
        return MFMA ? mfma_bin(kbase, src & 31u, src >> 5, m) : wbase + (uint32_t)m * T + src;
    }, P);
    OUT *pg = static_cast<OUT *>(a.pgram);
    Peaks L;
    peaks_init(L);
#pragma unroll
    for (int m = 0; m < M; ++m) {
        const OUT pw = (OUT)(P[m] * scale);
        if (kf[m] < a.nf) {
            if (pg) pg[(uint64_t)obj * a.nf + kf[m]] = pw;
            peaks_add(L, a.top_k, (double)pw, kf[m]);
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
// Slice sums of thread tid: [(q M + m) T + tid] (MFMA: the matrix-core
// result layout, m = result entry v).
template <bool SS_SUM, bool GEN, typename OUT, int T, int M, bool MFMA = false>
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
    split_epilogue<SS_SUM, GEN, OUT, T, M, NS, MFMA>(a, acc, obj, chunk, s_pp, s_pi);
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
// v_mfma_f32_32x32x16_f16 with FP32 accumulation (each instruction adds the
// exact sum of 8 products to the accumulator and rounds, twice: measured on
// the device).  Every factor x is split into an FP16 head h = fp16(x) and a
// tail l = fp16(x - h); head x head and the two cross products carry ~22 bits
// (the dropped tail x tail term is 2^-22 of the product); each complex sum's
// head x head and cross products accumulate in FP32 (apart, or in one sum for
// Z2 of the standard variant), unscaled in FP64.  The head in the MFMA
// operand must be the very value the tail was taken from (see split16x2).
// The row factor of row a is u = u0 R: u0 = e^{2 pi i f0 t} with f0 the
// item's first bin, one per observation and item (FP64 phase and sin/cos,
// staged in LDS), and R = e^{2 pi i 32 a df t}, one per observation and row
// of the object (k_colfac, FP64 phase and sin/cos); their FP32 product costs
// four FP32 operations per row and observation, and |u| = 1 to a few FP32
// ulps, which CC = (N + C2) / 2 assumes.  Row factors are unscaled (|u| <= 1:
// an FP16-subnormal head or tail costs at most 2^-24 absolute, whichever way
// the matrix core treats it); column factors are scaled by 2^10.
// The column factor depends only on the object: k_colfac writes its operand
// fragments once per 16 observations, and y is scaled per object by a power
// of two to max |y| in [0.5, 1) (FP16 range; undone exactly in the power).
// Observations past the object's end have zero row and column factors, so
// partial blocks need no masking.
// Operand layout of 32x32x16 (checked on the device): lane l holds row (A) or
// column (B) l % 32 and observations 8 (l / 32) .. +7 of the block; the
// result holds column l % 32, rows 8 (v / 4) + 4 (l / 32) + v % 4, v < 16.
// ---------------------------------------------------------------------------
typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));

constexpr float kFacScale = 1024.0f;  // 2^10 on the column factors

// Heads and tails of two factors x: h = fp16(x), l = fp16(x - h) (x - h is
// exact in FP32).  The head must be one value.  When x is a product a b, the compiler may
// round the exact product to FP16 in one step for the tail's head
// (v_fma_mixlo_f16 h, a, b, 0) yet convert the FP32-rounded x for the packed
// MFMA operand (v_cvt_pk_f16_f32); the two roundings differ by one FP16 ulp
// in ~1 of 2^13 factors, which are then wrong by 2^-11 (seen as 4e-3 power
// errors).  Here the two heads of a pair of factors come from one packed
// conversion of their FP32 values (v_cvt_pk_f16_f32), and the tails are
// taken against those very heads (the compiler may form x - h from the exact
// product; that only makes h + l closer to it).  tests/test_split_rounding.py
// checks the built kernels for the one-step product-to-FP16 form.
typedef _Float16 half2_t __attribute__((ext_vector_type(2)));
typedef float float2_t __attribute__((ext_vector_type(2)));

__device__ __forceinline__ void split16x2(float x0, float x1, half2_t &h, half2_t &l) {
    h = __builtin_convertvector((float2_t){x0, x1}, half2_t);
    const float2_t hf = __builtin_convertvector(h, float2_t);
    l = __builtin_convertvector((float2_t){x0 - hf.x, x1 - hf.y}, half2_t);
}

// First column-factor slot of an object: the slots of object k are
// [floor(beg/16) + k, floor(beg/16) + k + ceil(N/16)), which never reach the
// next object's floor((beg + N)/16) + k + 1.
__device__ __forceinline__ uint64_t colfac_slot(uint64_t beg, uint32_t obj) { return (beg >> 4) + obj; }

// Heads and tails of two FP64 factors: h = fp16(fp32(x)), l = fp16(x - h),
// the remainder x - h exact in FP64 (~22 bits of x in h + l).
__device__ __forceinline__ void split64x2(double x0, double x1, half2_t &h, half2_t &l) {
    h = __builtin_convertvector((float2_t){(float)x0, (float)x1}, half2_t);
    l = __builtin_convertvector((float2_t){(float)(x0 - (double)h.x), (float)(x1 - (double)h.y)}, half2_t);
}

// Wavefronts (object, y) with y < gridDim.y write the column-factor fragments
// of 16-observation blocks y, y + gridDim.y, ... of the object, each factor
// as head and tail of the real and imaginary parts (a product that needs -f_i
// takes the row factor's negated fragment instead):
//   standard:    y v (4 fragments), w (4)
//   generalized: (w/W) v (4), (w/W) y_c v (4), (w/W) w (4)
// with v = e^{2 pi i b df t} (b < 32 the column) and w = v^2, from the
// caller's FP64 (or FP32) t, y, dy in FP64 arithmetic: FP64 phase and
// sin/cos, w/W = 1 / (dy^2 W) the normalised weight, y_c = y - ybar; y (or
// (w/W) y_c) scaled per object by a power of two ys, and w/W by a power of
// two ws, to a maximum in [0.5, 1), then all by 2^10 (FP16 range; the scales
// are undone exactly in the power; wavefront y = 0 stores them); and the
// row-step fragments (e^{2 pi i 32 a df t}, row a = l % 32).  Observations
// past the object's end have zero factors.
template <bool GEN, typename IN>
__global__ __launch_bounds__(64) void k_colfac(SweepArgs a) {
    constexpr int NF = GEN ? kColfacFragsGen : kColfacFragsStd;
    const uint32_t k = blockIdx.x, l = threadIdx.x;
    const uint64_t beg = a.offsets[k], end = a.offsets[k + 1];
    const uint32_t n = (uint32_t)(end - beg);
    const IN *y = static_cast<const IN *>(a.y);
    const IN *dy = static_cast<const IN *>(a.dy);
    double iW = 1.0, ybar = 0.0;
    if constexpr (GEN) {
        iW = 1.0 / a.obj_gen[kGen * k + 4];
        ybar = a.obj_gen[kGen * k + 3];
    }
    auto weight = [&](uint64_t i) -> double {  // w / W (1 for the standard variant)
        if constexpr (GEN) {
            const double d = (double)dy[i];
            return iW / (d * d);
        } else {
            return 1.0;
        }
    };
    auto yfac = [&](uint64_t i) -> double {  // y, or (w / W) (y - ybar)
        if constexpr (GEN) return weight(i) * ((double)y[i] - ybar);
        else return (double)y[i];
    };
    double my = 0.0, mw = 0.0;
    for (uint64_t i = beg + l; i < end; i += 64) {
        my = fmax(my, fabs(yfac(i)));
        mw = fmax(mw, weight(i));
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        my = fmax(my, __shfl_xor(my, off, 64));
        mw = fmax(mw, __shfl_xor(mw, off, 64));
    }
    int ey = 0, ew = 0;
    (void)frexp(my, &ey);  // my = f 2^ey, f in [0.5, 1)
    (void)frexp(mw, &ew);
    const double ys = my > 0.0 ? ldexp(1.0, -ey) : 1.0;
    const double ws = mw > 0.0 ? ldexp(1.0, -ew) : 1.0;
    if (blockIdx.y == 0 && l == 0) {
        a.obj_ys[2 * k] = (float)ys;
        a.obj_ys[2 * k + 1] = (float)ws;
    }
    const double sy = ys * (double)kFacScale, sw = ws * (double)kFacScale;  // exact: powers of two
    const double fb = (double)(l & 31u) * a.df_rev;
    const double fr = (double)(32u * (l & 31u)) * a.df_rev;  // row step of row l % 32
    uint4 *dst = a.colfac + colfac_slot(beg, k) * (uint64_t)(NF * 64) + l;
    uint4 *rdst = a.rowfac + colfac_slot(beg, k) * (uint64_t)(kRowfacFrags * 64) + l;
    for (uint32_t q = blockIdx.y; q * 16u < n; q += gridDim.y) {
        // row-step phasors R = e^{2 pi i 32 (l % 32) df t} of observations
        // 8 (l / 32) .. + 7, in FP64 rounded to FP32 (zero past the end)
#pragma unroll
        for (int p = 0; p < kRowfacFrags; ++p) {
            float r4[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int e = 0; e < 2; ++e) {
                const uint32_t j = q * 16u + 8u * (l >> 5) + (uint32_t)(2 * p + e);
                if (j < n) {
                    double c, sn;
                    phasor64(fr, a.rec_t[beg + j], c, sn);
                    r4[2 * e] = (float)c;
                    r4[2 * e + 1] = (float)sn;
                }
            }
            uint4 u;
            __builtin_memcpy(&u, r4, sizeof(u));
            rdst[(uint64_t)(q * kRowfacFrags + p) * 64] = u;
        }
        half8_t f[NF];
#pragma unroll
        for (int v = 0; v < 8; v += 2) {
            double x[NF / 2][2];  // the factors of observations v, v + 1: real, imaginary per sum
#pragma unroll
            for (int e = 0; e < 2; ++e) {
                const uint32_t j = q * 16u + 8u * (l >> 5) + (uint32_t)(v + e);
                double cb = 0.0, sb = 0.0, yv = 0.0, wv = 0.0;
                if (j < n) {
                    phasor64(fb, a.rec_t[beg + j], cb, sb);
                    yv = yfac(beg + j) * sy;
                    wv = weight(beg + j) * sw;
                }
                const double wr = cb * cb - sb * sb, wi = 2.0 * cb * sb;  // v^2
                if constexpr (GEN) {
                    x[0][e] = wv * cb;
                    x[1][e] = wv * sb;
                    x[2][e] = yv * cb;
                    x[3][e] = yv * sb;
                    x[4][e] = wv * wr;
                    x[5][e] = wv * wi;
                } else {
                    x[0][e] = yv * cb;
                    x[1][e] = yv * sb;
                    x[2][e] = wv * wr;
                    x[3][e] = wv * wi;
                }
            }
#pragma unroll
            for (int p = 0; p < NF / 4; ++p) {  // each complex factor: r (h, l), i (h, l)
                half2_t h, t;
                split64x2(x[2 * p][0], x[2 * p][1], h, t);
                f[4 * p + 0][v] = h.x; f[4 * p + 0][v + 1] = h.y;
                f[4 * p + 1][v] = t.x; f[4 * p + 1][v + 1] = t.y;
                split64x2(x[2 * p + 1][0], x[2 * p + 1][1], h, t);
                f[4 * p + 2][v] = h.x; f[4 * p + 2][v + 1] = h.y;
                f[4 * p + 3][v] = t.x; f[4 * p + 3][v + 1] = t.y;
            }
        }
#pragma unroll
        for (int i = 0; i < NF; ++i) {
            uint4 u;
            __builtin_memcpy(&u, &f[i], sizeof(u));
            dst[(uint64_t)(q * NF + i) * 64] = u;
        }
    }
}

// Row factors of a 16-observation block for lane (row rc, half kh): u = u0 R
// with u0 = e^{2 pi i f0 t} (f0 the item's first bin; LDS, from s_u0[j0])
// and R = e^{2 pi i 32 rc df t} (the row-step fragments rr), and u^2, as FP16
// heads and tails.  Both phasors come from FP64 phases and FP64 sin/cos
// rounded to FP32, so |u| = 1 to a few FP32 ulps and u^2 needs no
// normalisation; the factors are unscaled (|u| <= 1).
struct RowFrags {
    half8_t ur_h, ur_l, ui_h, ui_l, qr_h, qr_l, qi_h, qi_l;
    half8_t nui_h, nui_l, nqi_h, nqi_l;  // -u_i, -(u^2)_i: the sign bits flipped
};
__device__ __forceinline__ half8_t neg8(half8_t x) {
    uint4 u;
    __builtin_memcpy(&u, &x, sizeof(u));
    u.x ^= 0x80008000u;
    u.y ^= 0x80008000u;
    u.z ^= 0x80008000u;
    u.w ^= 0x80008000u;
    __builtin_memcpy(&x, &u, sizeof(u));
    return x;
}
__device__ __forceinline__ void row_factors(const float2 *s_u0, uint32_t j0, uint32_t kh,
                                            const uint4 (&rr)[kRowfacFrags], RowFrags &R) {
#pragma unroll
    for (int v = 0; v < 8; v += 2) {
        float x[4][2];  // u_r, u_i, (u^2)_r, (u^2)_i
        const uint4 r = rr[v >> 1];
        const float rx[2] = {__uint_as_float(r.x), __uint_as_float(r.z)};
        const float ry[2] = {__uint_as_float(r.y), __uint_as_float(r.w)};
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            const float2 b = s_u0[j0 + 8u * kh + (uint32_t)(v + e)];  // < T: j0 <= T - 16
            const float ur = fmaf(b.x, rx[e], -(b.y * ry[e]));
            const float ui = fmaf(b.x, ry[e], b.y * rx[e]);
            x[0][e] = ur;
            x[1][e] = ui;
            x[2][e] = fmaf(ur, ur, -(ui * ui));
            x[3][e] = (ur + ur) * ui;
        }
        half2_t h, t;
        split16x2(x[0][0], x[0][1], h, t);
        R.ur_h[v] = h.x; R.ur_h[v + 1] = h.y;
        R.ur_l[v] = t.x; R.ur_l[v + 1] = t.y;
        split16x2(x[1][0], x[1][1], h, t);
        R.ui_h[v] = h.x; R.ui_h[v + 1] = h.y;
        R.ui_l[v] = t.x; R.ui_l[v + 1] = t.y;
        split16x2(x[2][0], x[2][1], h, t);
        R.qr_h[v] = h.x; R.qr_h[v + 1] = h.y;
        R.qr_l[v] = t.x; R.qr_l[v + 1] = t.y;
        split16x2(x[3][0], x[3][1], h, t);
        R.qi_h[v] = h.x; R.qi_h[v + 1] = h.y;
        R.qi_l[v] = t.x; R.qi_l[v + 1] = t.y;
    }
    R.nui_h = neg8(R.ui_h);
    R.nui_l = neg8(R.ui_l);
    R.nqi_h = neg8(R.qi_h);
    R.nqi_l = neg8(R.qi_l);
}

// Loads the row-step fragments of one 16-observation block.
__device__ __forceinline__ void load_rows(const uint4 *rf, uint4 (&rr)[kRowfacFrags]) {
#pragma unroll
    for (int i = 0; i < kRowfacFrags; ++i) rr[i] = rf[i * 64];
}

// e^{2 pi i f0 t} of observation b0 + l of the current stage into s_u0[l]
// (zero past cnt, the object's end: zero row factors).
__device__ __forceinline__ void stage_u0(const SweepArgs &a, uint64_t b0, uint32_t cnt, double f0, float2 *s_u0) {
    const uint32_t l = threadIdx.x;
    float2 u0 = make_float2(0.0f, 0.0f);
    if (l < cnt) {
        double c, sn;
        phasor64(f0, a.rec_t[b0 + l], c, sn);
        u0 = make_float2((float)c, (float)sn);
    }
    s_u0[l] = u0;
}

// Loads the column-factor fragments of one 16-observation block.
template <int NF>
__device__ __forceinline__ void load_cols(const uint4 *cf, half8_t (&col)[NF]) {
#pragma unroll
    for (int i = 0; i < NF; ++i) {
        const uint4 u = cf[i * 64];
        __builtin_memcpy(&col[i], &u, sizeof(u));
    }
}

#define LS_MFMA(ACC, A, B) ACC = __builtin_amdgcn_mfma_f32_32x32x16_f16(A, B, ACC, 0, 0, 0)

// Emits a lane's 16 bins of a matrix-core work item.  power(v, ill) gives
// result entry v's power (before the object's scale) and whether its
// condition estimate exceeds the bound; the other bins are stored and enter
// the peaks at once (kf ascends in v), the flagged ones (rare) are computed
// again in FP64 afterwards, one at a time with the whole wavefront, and merged
// with the index tie-break, so that no array of 16 powers stays live.
template <bool GEN, typename OUT, typename PowerOf>
__device__ __forceinline__ void mfma_emit(const SweepArgs &a, uint32_t obj, uint32_t chunk, uint32_t kbase,
                                          PowerOf power, double (*s_pp)[kPeaks], uint32_t (*s_pi)[kPeaks]) {
    const uint32_t l = threadIdx.x, rc = l & 31u, kh = l >> 5;
    const double scale = a.obj_scale[obj];
    OUT *pg = static_cast<OUT *>(a.pgram);
    const uint64_t row = (uint64_t)obj * a.nf;
    Peaks L;
    peaks_init(L);
    uint32_t flags = 0u;
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        const uint32_t kf = mfma_bin(kbase, rc, kh, v);
        bool ill;
        const double P = power(v, ill);
        if (kf < a.nf) {
            if (ill) {
                flags |= 1u << v;
            } else {
                const OUT pw = (OUT)(P * scale);
                if (pg) pg[row + kf] = pw;
                peaks_add(L, a.top_k, (double)pw, kf);
            }
        }
    }
    uint64_t lanes = __ballot(flags != 0u);
    while (lanes) {
        const uint32_t src = (uint32_t)__builtin_ctzll(lanes);
        const uint32_t f = (uint32_t)__shfl((int)flags, (int)src, 64);
        const uint32_t kf = mfma_bin(kbase, src & 31u, src >> 5, __builtin_ctz(f));
        const double pe = exact_power<GEN, OUT>(a, obj, kf);
        if (l == src) {
            const OUT pw = (OUT)(pe * scale);
            if (pg) pg[row + kf] = pw;
            peaks_merge(L, a.top_k, (double)pw, kf);
            flags &= flags - 1u;
        }
        lanes = __ballot(flags != 0u);
    }
    if (a.top_k) write_partial<OUT, 64>(a, L, obj, chunk, s_pp, s_pi);
}


// Stores a lane's 16 powers (scaled by the object's scale) and merges the
// item's peaks; kf ascends in v.
template <typename OUT>
__device__ __forceinline__ void mfma_store(const SweepArgs &a, uint32_t obj, uint32_t chunk, const uint32_t (&kf)[16],
                                           const double (&P)[16], double scale, double (*s_pp)[kPeaks],
                                           uint32_t (*s_pi)[kPeaks]) {
    OUT *pg = static_cast<OUT *>(a.pgram);
    Peaks L;
    peaks_init(L);
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        const OUT pw = (OUT)(P[v] * scale);
        if (kf[v] < a.nf) {
            if (pg) pg[(uint64_t)obj * a.nf + kf[v]] = pw;
            peaks_add(L, a.top_k, (double)pw, kf[v]);
        }
    }
    if (a.top_k) write_partial<OUT, 64>(a, L, obj, chunk, s_pp, s_pi);
}

// Two waves per SIMD: six FP32 accumulators (96 registers; Z1 keeps its
// cross products apart, where merging them doubles the rounding error of the
// power, while Z2 takes all three products in one) leave room for a second
// workgroup's registers, and the fold of the 256-observation partials stays
// in FP32 (16 KiB of LDS, 8 workgroups per CU).
template <typename OUT>
__global__ __launch_bounds__(64, 2) void k_sweep_mfma(SweepArgs a) {
    constexpr int T = 64, M = 16;  // 64 threads, 1024 bins per item
    constexpr uint32_t kFoldObs = 256;  // MFMA accumulation length
    __shared__ float2 s_u0[T];
    __shared__ float s_acc[4][16][T];  // sums of the 256-observation partials XC, XS, C2, S2
    __shared__ double s_pp[1][kPeaks];
    __shared__ uint32_t s_pi[1][kPeaks];
    __shared__ uint64_t s_item;
    const uint32_t l = threadIdx.x, rc = l & 31u, kh = l >> 5;
    constexpr double kUnscale = 1.0 / (double)kFacScale;  // column factors scaled, rows not
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
        const uint32_t kbase = chunk * (uint32_t)(T * M);
        if (oend - obeg <= kSmallObs && a.slices == 1) {  // the whole item in FP64
            double P[16];
            uint32_t kf[16];
#pragma unroll
            for (int v = 0; v < 16; ++v) kf[v] = mfma_bin(kbase, rc, kh, v);
            small_object_powers<false, OUT>(a, obj, kbase, rc, kh, P);
            mfma_store<OUT>(a, obj, chunk, kf, P, a.obj_scale[obj], s_pp, s_pi);
            continue;
        }
        constexpr int NF = kColfacFragsStd;
        const float ys = a.obj_ys[2 * obj], ws = a.obj_ys[2 * obj + 1];
        const double f0 = a.fmin_rev + (double)kbase * a.df_rev;  // the item's first bin
        const uint64_t slot = colfac_slot(obeg, obj) + ((beg - obeg) >> 4);
        const uint4 *cf = a.colfac + slot * (uint64_t)(NF * 64) + l;
        const uint4 *rf = a.rowfac + slot * (uint64_t)(kRowfacFrags * 64) + l;
        f32x16_t z1r = {}, z1r_x = {}, z1i = {}, z1i_x = {};  // Z1: head x head, cross products
        f32x16_t z2r = {}, z2i = {};
        // after every kFoldObs observations the accumulators move into FP32
        // sums in LDS (each lane its own entries, so no barrier); entry (q, v)
        // of this lane is acc_l[(16 q + v) T]
        float *const acc_l = &s_acc[0][0][l];
        bool folded = false;
        auto fold = [&]() {
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                float *e = acc_l + v * T;
                e[0 * 16 * T] = (folded ? e[0 * 16 * T] : 0.0f) + (z1r[v] + z1r_x[v]);
                e[1 * 16 * T] = (folded ? e[1 * 16 * T] : 0.0f) + (z1i[v] + z1i_x[v]);
                e[2 * 16 * T] = (folded ? e[2 * 16 * T] : 0.0f) + z2r[v];
                e[3 * 16 * T] = (folded ? e[3 * 16 * T] : 0.0f) + z2i[v];
            }
            z1r = f32x16_t{};
            z1r_x = f32x16_t{};
            z1i = f32x16_t{};
            z1i_x = f32x16_t{};
            z2r = f32x16_t{};
            z2i = f32x16_t{};
            folded = true;
        };
        // the row-step fragments are loaded one block ahead (their latency
        // would otherwise stall the row factors of every block)
        uint4 rr[kRowfacFrags];
        load_rows(rf, rr);
        for (uint64_t g0 = beg; g0 < end; g0 += kFoldObs) {
        const uint64_t gend = (end - g0) < kFoldObs ? end : g0 + kFoldObs;
        for (uint64_t b0 = g0; b0 < gend; b0 += T) {
            const uint32_t cnt = (uint32_t)((gend - b0) < (uint64_t)T ? (gend - b0) : (uint64_t)T);
            __syncthreads();  // also orders the previous item's peak merge before this one's
            stage_u0(a, b0, cnt, f0, s_u0);
            __syncthreads();
            for (uint32_t j0 = 0; j0 < cnt; j0 += 16, cf += NF * 64, rf += kRowfacFrags * 64) {
                half8_t col[NF];
                load_cols<NF>(cf, col);
                RowFrags R;
                row_factors(s_u0, j0, kh, rr, R);
                // the next block's, loaded while this one's MFMAs run (one slot
                // past an object's or slice's last block is still inside the
                // table: at most slot n/16 + K of n/16 + K + 1)
                load_rows(rf + kRowfacFrags * 64, rr);
                // Re Z1 = u_r (y v)_r - u_i (y v)_i,  Im Z1 = u_r (y v)_i + u_i (y v)_r
                LS_MFMA(z1r, R.ur_h, col[0]);
                LS_MFMA(z1r_x, R.ur_h, col[1]);
                LS_MFMA(z1r_x, R.ur_l, col[0]);
                LS_MFMA(z1r, R.nui_h, col[2]);
                LS_MFMA(z1r_x, R.nui_h, col[3]);
                LS_MFMA(z1r_x, R.nui_l, col[2]);
                LS_MFMA(z1i, R.ur_h, col[2]);
                LS_MFMA(z1i_x, R.ur_h, col[3]);
                LS_MFMA(z1i_x, R.ur_l, col[2]);
                LS_MFMA(z1i, R.ui_h, col[0]);
                LS_MFMA(z1i_x, R.ui_h, col[1]);
                LS_MFMA(z1i_x, R.ui_l, col[0]);
                // the same for Z2 with u^2 and w
                LS_MFMA(z2r, R.qr_h, col[4]);
                LS_MFMA(z2r, R.qr_h, col[5]);
                LS_MFMA(z2r, R.qr_l, col[4]);
                LS_MFMA(z2r, R.nqi_h, col[6]);
                LS_MFMA(z2r, R.nqi_h, col[7]);
                LS_MFMA(z2r, R.nqi_l, col[6]);
                LS_MFMA(z2i, R.qr_h, col[6]);
                LS_MFMA(z2i, R.qr_h, col[7]);
                LS_MFMA(z2i, R.qr_l, col[6]);
                LS_MFMA(z2i, R.qi_h, col[4]);
                LS_MFMA(z2i, R.qi_h, col[5]);
                LS_MFMA(z2i, R.qi_l, col[4]);
            }
        }
        if (gend < end) fold();
        }
        // the item's FP64 sums of result entry v
        // (head x head + cross products in FP32: one rounding relative to the
        // sum itself)
        auto sums = [&](int v, double &XC, double &XS, double &C2, double &S2) {
            XC = (double)(z1r[v] + z1r_x[v]);
            XS = (double)(z1i[v] + z1i_x[v]);
            C2 = (double)z2r[v];
            S2 = (double)z2i[v];
            if (folded) {
                const float *e = acc_l + v * T;
                XC += (double)e[0 * 16 * T];
                XS += (double)e[1 * 16 * T];
                C2 += (double)e[2 * 16 * T];
                S2 += (double)e[3 * 16 * T];
            }
        };
        const double uy = kUnscale / (double)ys, uw = kUnscale / (double)ws;  // powers of two
        if (a.slices > 1) {
            // the slice's FP64 sums XC, XS, CC, CS of result entry v at
            // [(q M + v) T + l], the matrix-core layout k_combine_slices reads
            double *dst = a.part_sums + (((uint64_t)obj * a.chunks + chunk) * a.slices + slice) *
                                            (uint64_t)(4 * M * T) + l;
            const double ns = (double)(end - beg);
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                double XC, XS, C2, S2;
                sums(v, XC, XS, C2, S2);
                dst[(0 * M + v) * T] = XC * uy;
                dst[(1 * M + v) * T] = XS * uy;
                dst[(2 * M + v) * T] = 0.5 * (ns + C2 * uw);
                dst[(3 * M + v) * T] = 0.5 * (S2 * uw);
            }
            continue;
        }
        // std_power with the scales folded in: CC = (N + uw C2) / 2, CS = uw S2 / 2,
        // SS = N - CC; the unscaled XC, XS enter the numerator, scaled by uy^2
        const double nobs = (double)(oend - obeg), hn = 0.5 * nobs, hw = 0.5 * uw;
        const double dmin = nobs * nobs / (4.0 * kCondMax), hy2 = 0.5 * uy * uy;
        mfma_emit<false, OUT>(a, obj, chunk, kbase, [&](int v, bool &ill) {
            double XC, XS, C2, S2;
            sums(v, XC, XS, C2, S2);
            const double CC = fma(C2, hw, hn), CS = S2 * hw, SS = nobs - CC;
            const double D = fma(CC, SS, -(CS * CS));
            ill = !(D >= dmin);  // (CC + SS)^2 / (4 D) > kCondMax, D <= 0 or NaN
            const double num = fma(SS, XC * XC, fma(CC, XS * XS, -2.0 * CS * (XC * XS)));
            return fdiv(num * hy2, D);
        }, s_pp, s_pi);
    }
}


// Generalized variant of the matrix-core tile.  With q = sqrt(w / W) (the
// records' normalised weights) and y_c = y - ybar, three complex sums per bin:
//   Zw = sum q^2 e^{i phi},  Zy = sum q (q y_c) e^{i phi},  Z2 = sum q^2 e^{2 i phi}
// give the seven weighted sums of the generalized power: C, S = Zw;
// YC, YS = Zy; CC, SS = (Wf +- Re Z2) / 2; CS = Im Z2 / 2, with Wf = sum q^2.
// 36 MFMAs per 16 observations.
// All three sums accumulate in FP32 (head x head and cross products apart)
// across 256 observations and then fold into sums in LDS: Zy, which sets the
// precision of weak bins (P ~ 1e-3 P_max), into FP64; the weight-only Zw and
// Z2 into FP32 (they feed the centred CC - C^2, SS - S^2, CS - C S, which
// cancel at aliasing frequencies, and those bins exceed the condition bound
// and are recomputed in FP64).  Each MFMA adds exact 8-product sums to its
// accumulator, so restarting Zy per 16-observation block gains nothing.
// One wavefront per SIMD (more than 256 registers per lane); 32 KiB of LDS.
template <typename OUT>
__global__ __launch_bounds__(64) void k_sweep_mfma_gen(SweepArgs a) {
    constexpr int T = 64, M = 16;
    constexpr uint32_t kFoldObs = 256;
    __shared__ float2 s_u0[T];
    __shared__ float s_accw[4][16][T];   // FP32 Re/Im of Zw and Z2 of a lane's 16 bins
    __shared__ double s_accy[2][16][T];  // FP64 Re/Im of Zy
    __shared__ double s_pp[1][kPeaks];
    __shared__ uint32_t s_pi[1][kPeaks];
    __shared__ uint64_t s_item;
    const uint32_t l = threadIdx.x, rc = l & 31u, kh = l >> 5;
    constexpr double kUnscale = 1.0 / (double)kFacScale;  // column factors scaled, rows not
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
        const uint32_t kbase = chunk * (uint32_t)(T * M);
        if (oend - obeg <= kSmallObs && a.slices == 1) {  // the whole item in FP64
            double P[16];
            uint32_t kf[16];
#pragma unroll
            for (int v = 0; v < 16; ++v) kf[v] = mfma_bin(kbase, rc, kh, v);
            small_object_powers<true, OUT>(a, obj, kbase, rc, kh, P);
            mfma_store<OUT>(a, obj, chunk, kf, P, a.obj_scale[obj], s_pp, s_pi);
            continue;
        }
        constexpr int NF = kColfacFragsGen;
        const float ys = a.obj_ys[2 * obj], ws = a.obj_ys[2 * obj + 1];
        const double f0 = a.fmin_rev + (double)kbase * a.df_rev;  // the item's first bin
        const uint64_t slot = colfac_slot(obeg, obj) + ((beg - obeg) >> 4);
        const uint4 *cf = a.colfac + slot * (uint64_t)(NF * 64) + l;
        const uint4 *rf = a.rowfac + slot * (uint64_t)(kRowfacFrags * 64) + l;
        // {Re hh, Re x, Im hh, Im x} since the last fold
        f32x16_t zw[4] = {}, z2[4] = {}, zy[4] = {};
        // entry (q, v) of this lane's LDS sums: acc_l[(16 q + v) T], acy_l likewise
        float *const acc_l = &s_accw[0][0][l];
        double *const acy_l = &s_accy[0][0][l];
        bool folded = false;
        auto fold = [&]() {
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                float *e = acc_l + v * T;
                e[0 * 16 * T] = (folded ? e[0 * 16 * T] : 0.0f) + (zw[0][v] + zw[1][v]);
                e[1 * 16 * T] = (folded ? e[1 * 16 * T] : 0.0f) + (zw[2][v] + zw[3][v]);
                e[2 * 16 * T] = (folded ? e[2 * 16 * T] : 0.0f) + (z2[0][v] + z2[1][v]);
                e[3 * 16 * T] = (folded ? e[3 * 16 * T] : 0.0f) + (z2[2][v] + z2[3][v]);
                double *d = acy_l + v * T;
                d[0] = (folded ? d[0] : 0.0) + ((double)zy[0][v] + (double)zy[1][v]);
                d[16 * T] = (folded ? d[16 * T] : 0.0) + ((double)zy[2][v] + (double)zy[3][v]);
            }
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                zw[i] = f32x16_t{};
                z2[i] = f32x16_t{};
                zy[i] = f32x16_t{};
            }
            folded = true;
        };
        // the row-step fragments are loaded one block ahead (their latency
        // would otherwise stall the row factors of every block)
        uint4 rr[kRowfacFrags];
        load_rows(rf, rr);
        for (uint64_t g0 = beg; g0 < end; g0 += kFoldObs) {
        const uint64_t gend = (end - g0) < kFoldObs ? end : g0 + kFoldObs;
        for (uint64_t b0 = g0; b0 < gend; b0 += T) {
            const uint32_t cnt = (uint32_t)((gend - b0) < (uint64_t)T ? (gend - b0) : (uint64_t)T);
            __syncthreads();  // also orders the previous item's peak merge before this one's
            stage_u0(a, b0, cnt, f0, s_u0);
            __syncthreads();
            for (uint32_t j0 = 0; j0 < cnt; j0 += 16, cf += NF * 64, rf += kRowfacFrags * 64) {
                half8_t col[NF];
                load_cols<NF>(cf, col);
                RowFrags R;
                row_factors(s_u0, j0, kh, rr, R);
                // the next block's, loaded while this one's MFMAs run (one slot
                // past an object's or slice's last block is still inside the
                // table: at most slot n/16 + K of n/16 + K + 1)
                load_rows(rf + kRowfacFrags * 64, rr);
                // Re Z = u_r f_r - u_i f_i,  Im Z = u_r f_i + u_i f_r for the
                // column factors f = q^2 v (col 0..5) of Zw, q (q y) v (6..11)
                // of Zy, and, with u^2/|u|^2 for u, q^2 w (12..17) of Z2;
                // head x head ([0], [2]) and cross products ([1], [3]) apart.
                LS_MFMA(zw[0], R.ur_h, col[0]);
                LS_MFMA(zw[1], R.ur_h, col[1]);
                LS_MFMA(zw[1], R.ur_l, col[0]);
                LS_MFMA(zw[0], R.nui_h, col[2]);
                LS_MFMA(zw[1], R.nui_h, col[3]);
                LS_MFMA(zw[1], R.nui_l, col[2]);
                LS_MFMA(zw[2], R.ur_h, col[2]);
                LS_MFMA(zw[3], R.ur_h, col[3]);
                LS_MFMA(zw[3], R.ur_l, col[2]);
                LS_MFMA(zw[2], R.ui_h, col[0]);
                LS_MFMA(zw[3], R.ui_h, col[1]);
                LS_MFMA(zw[3], R.ui_l, col[0]);
                LS_MFMA(z2[0], R.qr_h, col[8]);
                LS_MFMA(z2[1], R.qr_h, col[9]);
                LS_MFMA(z2[1], R.qr_l, col[8]);
                LS_MFMA(z2[0], R.nqi_h, col[10]);
                LS_MFMA(z2[1], R.nqi_h, col[11]);
                LS_MFMA(z2[1], R.nqi_l, col[10]);
                LS_MFMA(z2[2], R.qr_h, col[10]);
                LS_MFMA(z2[3], R.qr_h, col[11]);
                LS_MFMA(z2[3], R.qr_l, col[10]);
                LS_MFMA(z2[2], R.qi_h, col[8]);
                LS_MFMA(z2[3], R.qi_h, col[9]);
                LS_MFMA(z2[3], R.qi_l, col[8]);
                LS_MFMA(zy[1], R.ur_h, col[5]);
                LS_MFMA(zy[1], R.ur_l, col[4]);
                LS_MFMA(zy[1], R.nui_h, col[7]);
                LS_MFMA(zy[1], R.nui_l, col[6]);
                LS_MFMA(zy[0], R.ur_h, col[4]);
                LS_MFMA(zy[0], R.nui_h, col[6]);
                LS_MFMA(zy[3], R.ur_h, col[7]);
                LS_MFMA(zy[3], R.ur_l, col[6]);
                LS_MFMA(zy[3], R.ui_h, col[5]);
                LS_MFMA(zy[3], R.ui_l, col[4]);
                LS_MFMA(zy[2], R.ur_h, col[6]);
                LS_MFMA(zy[2], R.ui_h, col[4]);
            }
        }
        if (gend < end) fold();
        }
        // the item's FP64 sums of result entry v: Re/Im Zw, Zy (y unscaled), Z2
        // (head x head + cross products in FP32: one rounding relative to the
        // sum itself)
        auto sums = [&](int v, double (&z)[6]) {
            z[0] = (double)(zw[0][v] + zw[1][v]);
            z[1] = (double)(zw[2][v] + zw[3][v]);
            z[2] = (double)(zy[0][v] + zy[1][v]);
            z[3] = (double)(zy[2][v] + zy[3][v]);
            z[4] = (double)(z2[0][v] + z2[1][v]);
            z[5] = (double)(z2[2][v] + z2[3][v]);
            if (folded) {
                const float *e = acc_l + v * T;
                z[0] += (double)e[0 * 16 * T];
                z[1] += (double)e[1 * 16 * T];
                z[4] += (double)e[2 * 16 * T];
                z[5] += (double)e[3 * 16 * T];
                const double *d = acy_l + v * T;
                z[2] += d[0];
                z[3] += d[16 * T];
            }
        };
        const double uy = kUnscale / (double)ys, uw = kUnscale / (double)ws;  // powers of two
        if (a.slices > 1) {
            // the slice's seven FP64 sums in the matrix-core layout
            // k_combine_slices reads, with the slice's weight sum for CC and SS
            double wsum = 0.0;  // of w / W, as the column factors
            const OUT *dy = static_cast<const OUT *>(a.dy);
            const double iWt = 1.0 / a.obj_gen[kGen * obj + 4];
            for (uint64_t i = beg + l; i < end; i += T) {
                const double d = (double)dy[i];
                wsum += iWt / (d * d);
            }
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) wsum += __shfl_xor(wsum, off, 64);
            double *dst = a.part_sums + (((uint64_t)obj * a.chunks + chunk) * a.slices + slice) *
                                            (uint64_t)(7 * M * T) + l;
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                double z[6];
                sums(v, z);
                dst[(0 * M + v) * T] = z[0] * uw;
                dst[(1 * M + v) * T] = z[1] * uw;
                dst[(2 * M + v) * T] = z[2] * uy;
                dst[(3 * M + v) * T] = z[3] * uy;
                dst[(4 * M + v) * T] = 0.5 * (wsum + z[4] * uw);
                dst[(5 * M + v) * T] = 0.5 * (wsum - z[4] * uw);
                dst[(6 * M + v) * T] = 0.5 * (z[5] * uw);
            }
            continue;
        }
        // Wf, Yb, YY over the FP64 weights w / W of the column factors
        const double Wf = a.obj_gen[kGen * obj + 5], Yb = a.obj_gen[kGen * obj + 6];
        const double YY = a.obj_gen[kGen * obj + 7], iW = 1.0 / Wf;
        const double sw = uw * iW, sy = uy * iW, hsw = 0.5 * sw;  // scales with the weight sum folded in
        mfma_emit<true, OUT>(a, obj, chunk, kbase, [&](int v, bool &ill) {
            double z[6];
            sums(v, z);
            const double h2 = z[4] * hsw;  // (Re Z2 / Wf) / 2
            return gen_power(z[0] * sw, z[1] * sw, 0.5 + h2, 0.5 - h2, z[5] * hsw, z[2] * sy, z[3] * sy, Yb, YY,
                             ill);
        }, s_pp, s_pi);
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
     k_combine_slices<true, false, float, 64, 16, true>, k_colfac<false, float>},
    {LSAMD_FORM_SPLIT32, 0, LSAMD_F64, 64, 16, k_sweep_mfma<double>,
     k_combine_slices<true, false, double, 64, 16, true>, k_colfac<false, double>},
    {LSAMD_FORM_SPLIT32, 1, LSAMD_F32, 64, 16, k_sweep_mfma_gen<float>,
     k_combine_slices<true, true, float, 64, 16, true>, k_colfac<true, float>},
    {LSAMD_FORM_SPLIT32, 1, LSAMD_F64, 64, 16, k_sweep_mfma_gen<double>,
     k_combine_slices<true, true, double, 64, 16, true>, k_colfac<true, double>},
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
