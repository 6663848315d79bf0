// Plain-form sweep kernels: the reference's own arithmetic, standard and
// generalized (source/main.cu:1566-1599, 1701-1777), for the FP32 and FP64
// builds, and their tile registry entries.

#include "ls_device.h"

namespace lsamd {

namespace {

// ---------------------------------------------------------------------------
// Plain form: the reference's arithmetic, standard variant
// (lombscarglecpuinnerloop, main.cu:1566-1599).
// ---------------------------------------------------------------------------
template <typename R, int T, int M>
__global__ __launch_bounds__(T) void k_sweep_plain_std(SweepArgs a) {
#pragma clang fp contract(off)
    __shared__ R s_t[T], s_y[T];
    __shared__ double s_pp[T / 64][kPeaks];
    __shared__ uint32_t s_pi[T / 64][kPeaks];
    __shared__ uint64_t s_item;
    for (uint64_t item = first_item(a, &s_item); item < a.total_items;
         item = next_item(a, &s_item, item)) {
        uint32_t obj, chunk;
        item_coords(a, item, obj, chunk);
        const uint64_t beg = a.offsets[obj], end = a.offsets[obj + 1];
        const R *gt = static_cast<const R *>(a.t);
        const R *gy = static_cast<const R *>(a.y);
        const uint32_t tid = threadIdx.x;
        const uint32_t kbase = chunk * (uint32_t)(T * M);
        const R fmin = (R)a.fmin_r, fstep = (R)a.fstep_r;
        R om[M], xc[M], xs[M], cc[M], ss[M], cs[M];
    #pragma unroll
        for (int m = 0; m < M; ++m) {
            om[m] = fmin + (fstep * (R)(kbase + m * T + tid));  // main.cu:1611
            xc[m] = xs[m] = cc[m] = ss[m] = cs[m] = (R)0;
        }
        for (uint64_t b0 = beg; b0 < end; b0 += T) {
            const uint32_t cnt = (uint32_t)((end - b0) < (uint64_t)T ? (end - b0) : (uint64_t)T);
            __syncthreads();
            if (tid < cnt) {
                s_t[tid] = gt[b0 + tid];
                s_y[tid] = gy[b0 + tid];
            }
            __syncthreads();
            for (uint32_t j = 0; j < cnt; ++j) {
                const R tj = s_t[j], yj = s_y[j];
    #pragma unroll
                for (int m = 0; m < M; ++m) {
                    R s, c;
                    sincos_r(om[m] * tj, &s, &c);
                    xc[m] = xc[m] + yj * c;
                    xs[m] = xs[m] + yj * s;
                    cc[m] = cc[m] + c * c;
                    ss[m] = ss[m] + s * s;
                    cs[m] = cs[m] + c * s;
                }
            }
        }
        const double scale = a.obj_scale[obj];
        R *pg = static_cast<R *>(a.pgram);
        Peaks L;
        peaks_init(L);
    #pragma unroll
        for (int m = 0; m < M; ++m) {
            const uint32_t kf = kbase + m * T + tid;
            const R omega = om[m];
            const R tau = (R)(atan2(2.0 * (double)cs[m], (double)(cc[m] - ss[m])) / (2.0 * (double)omega));
            R s_tau, c_tau;
            sincos_r(omega * tau, &s_tau, &c_tau);
            const R c_tau2 = c_tau * c_tau, s_tau2 = s_tau * s_tau;
            const R cs_tau = (R)(2.0 * (double)c_tau * (double)s_tau);
            const R p1 = c_tau * xc[m] + s_tau * xs[m];
            const R p2 = c_tau * xs[m] - s_tau * xc[m];
            const R q = (p1 * p1) / (c_tau2 * cc[m] + cs_tau * cs[m] + s_tau2 * ss[m]) +
                        (p2 * p2) / (c_tau2 * ss[m] - cs_tau * cs[m] + s_tau2 * cc[m]);
            R pw = (R)(0.5 * (double)q);
            pw = (R)((double)pw * scale);  // main.cu:1335 / 1350
            if (kf < a.nf) {
                if (pg) pg[(uint64_t)obj * a.nf + kf] = pw;
                peaks_add(L, a.top_k, (double)pw, kf);
            }
        }
        if (a.top_k) write_partial<R, T>(a, L, obj, chunk, s_pp, s_pi);
    }
}

// ---------------------------------------------------------------------------
// Plain form, generalized variant: two passes, tau then the tau-shifted sums
// (lombscarglecpuinnerloopAstroPy, main.cu:1701-1777).
// ---------------------------------------------------------------------------
template <typename R, int T, int M>
__global__ __launch_bounds__(T) void k_sweep_plain_gen(SweepArgs a) {
#pragma clang fp contract(off)
    __shared__ R s_t[T], s_y[T], s_w[T];
    __shared__ double s_pp[T / 64][kPeaks];
    __shared__ uint32_t s_pi[T / 64][kPeaks];
    __shared__ uint64_t s_item;
    for (uint64_t item = first_item(a, &s_item); item < a.total_items;
         item = next_item(a, &s_item, item)) {
        uint32_t obj, chunk;
        item_coords(a, item, obj, chunk);
        const uint64_t beg = a.offsets[obj], end = a.offsets[obj + 1];
        const R *gt = static_cast<const R *>(a.t);
        const R *gy = static_cast<const R *>(a.y);
        const R *gw = static_cast<const R *>(a.w);
        const uint32_t tid = threadIdx.x;
        const uint32_t kbase = chunk * (uint32_t)(T * M);
        const R fmin = (R)a.fmin_r, fstep = (R)a.fstep_r;
        R om[M], wsum[M], S[M], C[M], S2[M], C2[M], tau[M];
    #pragma unroll
        for (int m = 0; m < M; ++m) {
            om[m] = fmin + (fstep * (R)(kbase + m * T + tid));
            wsum[m] = S[m] = C[m] = S2[m] = C2[m] = (R)0;
        }
        // pass 1: tau
        for (uint64_t b0 = beg; b0 < end; b0 += T) {
            const uint32_t cnt = (uint32_t)((end - b0) < (uint64_t)T ? (end - b0) : (uint64_t)T);
            __syncthreads();
            if (tid < cnt) {
                s_t[tid] = gt[b0 + tid];
                s_w[tid] = gw[b0 + tid];
            }
            __syncthreads();
            for (uint32_t j = 0; j < cnt; ++j) {
                const R tj = s_t[j], w = s_w[j];
    #pragma unroll
                for (int m = 0; m < M; ++m) {
                    R s, c;
                    wsum[m] += w;
                    sincos_r(om[m] * tj, &s, &c);
                    S[m] += w * s;
                    C[m] += w * c;
                    S2[m] = (R)((double)S2[m] + 2.0 * (double)w * (double)s * (double)c);
                    C2[m] = (R)((double)C2[m] + ((double)w - 2.0 * (double)w * (double)s * (double)s));
                }
            }
        }
    #pragma unroll
        for (int m = 0; m < M; ++m) {
            S2[m] /= wsum[m];
            C2[m] /= wsum[m];
            S[m] /= wsum[m];
            C[m] /= wsum[m];
            S2[m] = (R)((double)S2[m] - 2.0 * (double)S[m] * (double)C[m]);
            C2[m] -= (C[m] * C[m] - S[m] * S[m]);
            tau[m] = (R)(0.5 * (double)atan2_r(S2[m], C2[m]) / (double)om[m]);
        }
        // pass 2: the tau-shifted sums
        R Y[M], YY[M], Ct[M], St[M], YC[M], YS[M], CC[M], SS[M];
    #pragma unroll
        for (int m = 0; m < M; ++m) Y[m] = YY[m] = Ct[m] = St[m] = YC[m] = YS[m] = CC[m] = SS[m] = (R)0;
        for (uint64_t b0 = beg; b0 < end; b0 += T) {
            const uint32_t cnt = (uint32_t)((end - b0) < (uint64_t)T ? (end - b0) : (uint64_t)T);
            __syncthreads();
            if (tid < cnt) {
                s_t[tid] = gt[b0 + tid];
                s_y[tid] = gy[b0 + tid];
                s_w[tid] = gw[b0 + tid];
            }
            __syncthreads();
            for (uint32_t j = 0; j < cnt; ++j) {
                const R tj = s_t[j], yj = s_y[j], w = s_w[j];
    #pragma unroll
                for (int m = 0; m < M; ++m) {
                    R s, c;
                    sincos_r(om[m] * (tj - tau[m]), &s, &c);
                    Y[m] += w * yj;
                    YY[m] += w * yj * yj;
                    Ct[m] += w * c;
                    St[m] += w * s;
                    YC[m] += w * yj * c;
                    YS[m] += w * yj * s;
                    CC[m] += w * c * c;
                    SS[m] += w * s * s;
                }
            }
        }
        R *pg = static_cast<R *>(a.pgram);
        Peaks L;
        peaks_init(L);
    #pragma unroll
        for (int m = 0; m < M; ++m) {
            const uint32_t kf = kbase + m * T + tid;
            const R ws = wsum[m];
            R y_ = Y[m] / ws, yy = YY[m] / ws, ct = Ct[m] / ws, st = St[m] / ws;
            R yc = YC[m] / ws, ys = YS[m] / ws, ccv = CC[m] / ws, ssv = SS[m] / ws;
            yc -= y_ * ct;
            ys -= y_ * st;
            ccv -= ct * ct;
            ssv -= st * st;
            yy -= y_ * y_;
            const R pw = (yc * yc / ccv + ys * ys / ssv) / yy;
            if (kf < a.nf) {
                if (pg) pg[(uint64_t)obj * a.nf + kf] = pw;
                peaks_add(L, a.top_k, (double)pw, kf);
            }
        }
        if (a.top_k) write_partial<R, T>(a, L, obj, chunk, s_pp, s_pi);
    }
}

}  // namespace

#define LS_PLAIN_STD(R, DT, TT, MM) \
    {LSAMD_FORM_PLAIN, 0, DT, TT, MM, k_sweep_plain_std<R, TT, MM>, nullptr}
#define LS_PLAIN_GEN(R, DT, TT, MM) \
    {LSAMD_FORM_PLAIN, 1, DT, TT, MM, k_sweep_plain_gen<R, TT, MM>, nullptr}

const TileEntry kPlainTiles[] = {
    LS_PLAIN_STD(float, LSAMD_F32, 256, 1), LS_PLAIN_STD(double, LSAMD_F64, 256, 1),
    LS_PLAIN_STD(float, LSAMD_F32, 256, 2), LS_PLAIN_STD(double, LSAMD_F64, 256, 2),
    LS_PLAIN_GEN(float, LSAMD_F32, 256, 1), LS_PLAIN_GEN(double, LSAMD_F64, 256, 1),
};
const int kPlainTileCount = (int)(sizeof(kPlainTiles) / sizeof(kPlainTiles[0]));

}  // namespace lsamd
