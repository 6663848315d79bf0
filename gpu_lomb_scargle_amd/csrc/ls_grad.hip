// Lomb-Scargle power at arbitrary frequencies with its derivatives, and
// Newton refinement of periodogram peaks (lsamd_power_grad,
// lsamd_refine_peaks).
//
// One wavefront per object.  Lanes stride over the object's observations,
// accumulate FP64 sums of (cos, sin) of omega*t and of 2*omega*t weighted by
// 1, t and t^2, and reduce them with a butterfly so every lane holds the
// totals.  The power is a closed-form function of nine raw sums,
//   generalized: W = sum u, A1 = sum u y, A2 = sum u y^2, C = sum u c,
//                S = sum u s, C2 = sum u cos 2wt, S2 = sum u sin 2wt,
//                YC = sum u y c, YS = sum u y s            (u = dy^-2)
//     P = (SS YC'^2 - 2 CS YC' YS' + CC YS'^2) / ((CC SS - CS^2) YY)
//     with the centred, W-normalised CC, SS, CS, YC', YS', YY;
//   standard:    P = (SS XC^2 - 2 CS XC XS + CC XS^2) / (2 (CC SS - CS^2))
//     with CC = (N + C2)/2, SS = (N - C2)/2, CS = S2/2, XC = YC, XS = YS,
//     times 2/sum (y - ybar)^2 when normalised,
// the tau-free form of the reference's power (main.cu:1566-1599 standard,
// main.cu:1701-1777 generalized).  That one function is evaluated twice
// through operator overloading: on jets (value, d/domega, d2/domega2) of
// the sums, whose omega-derivatives are the t- and t^2-weighted sums, and on
// duals carrying the nine partials dP/d(sum), which turn into per-
// observation gradients dP/dy_j and dP/d(dy_j) by the chain rule.
//
// Times are taken relative to the middle of each object's span: the power is
// invariant under t -> t + t0, and the t- and t^2-weighted sums stay small.
// Generalized magnitudes are centred on their weighted mean for the same
// reason (the generalized power is invariant under y -> y + k).

#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include "ls_internal.h"

namespace lsamd {

namespace {

// ---- forward-mode jets in omega (value, first, second derivative) ----------
struct J2 {
    double v, d, dd;
};
__device__ __forceinline__ J2 jc(double v) { return {v, 0.0, 0.0}; }
__device__ __forceinline__ J2 operator+(J2 a, J2 b) { return {a.v + b.v, a.d + b.d, a.dd + b.dd}; }
__device__ __forceinline__ J2 operator-(J2 a, J2 b) { return {a.v - b.v, a.d - b.d, a.dd - b.dd}; }
__device__ __forceinline__ J2 operator*(J2 a, J2 b) {
    return {a.v * b.v, a.d * b.v + a.v * b.d, a.dd * b.v + 2.0 * a.d * b.d + a.v * b.dd};
}
__device__ __forceinline__ J2 operator*(double s, J2 a) { return {s * a.v, s * a.d, s * a.dd}; }
__device__ __forceinline__ J2 operator+(double s, J2 a) { return {s + a.v, a.d, a.dd}; }
__device__ __forceinline__ J2 operator/(J2 a, J2 b) {
    // q = a/b: q' = (a' - q b')/b, q'' = (a'' - 2 q' b' - q b'')/b
    const double q = a.v / b.v;
    const double q1 = (a.d - q * b.d) / b.v;
    const double q2 = (a.dd - 2.0 * q1 * b.d - q * b.dd) / b.v;
    return {q, q1, q2};
}

// ---- forward-mode duals with N partials -----------------------------------
template <int N>
struct DN {
    double v;
    double g[N];
};
template <int N>
__device__ __forceinline__ DN<N> dseed(double v, int i) {
    DN<N> r;
    r.v = v;
#pragma unroll
    for (int k = 0; k < N; ++k) r.g[k] = (k == i) ? 1.0 : 0.0;
    return r;
}
template <int N>
__device__ __forceinline__ DN<N> dconst(double v) {
    DN<N> r;
    r.v = v;
#pragma unroll
    for (int k = 0; k < N; ++k) r.g[k] = 0.0;
    return r;
}
template <int N>
__device__ __forceinline__ DN<N> operator/(double s, const DN<N> &b) {
    DN<N> r;
    r.v = s / b.v;
#pragma unroll
    for (int k = 0; k < N; ++k) r.g[k] = -r.v * b.g[k] / b.v;
    return r;
}
template <int N>
__device__ __forceinline__ DN<N> operator+(const DN<N> &a, const DN<N> &b) {
    DN<N> r;
    r.v = a.v + b.v;
#pragma unroll
    for (int k = 0; k < N; ++k) r.g[k] = a.g[k] + b.g[k];
    return r;
}
template <int N>
__device__ __forceinline__ DN<N> operator-(const DN<N> &a, const DN<N> &b) {
    DN<N> r;
    r.v = a.v - b.v;
#pragma unroll
    for (int k = 0; k < N; ++k) r.g[k] = a.g[k] - b.g[k];
    return r;
}
template <int N>
__device__ __forceinline__ DN<N> operator*(const DN<N> &a, const DN<N> &b) {
    DN<N> r;
    r.v = a.v * b.v;
#pragma unroll
    for (int k = 0; k < N; ++k) r.g[k] = a.g[k] * b.v + a.v * b.g[k];
    return r;
}
template <int N>
__device__ __forceinline__ DN<N> operator*(double s, const DN<N> &a) {
    DN<N> r;
    r.v = s * a.v;
#pragma unroll
    for (int k = 0; k < N; ++k) r.g[k] = s * a.g[k];
    return r;
}
template <int N>
__device__ __forceinline__ DN<N> operator+(double s, const DN<N> &a) {
    DN<N> r = a;
    r.v = s + a.v;
    return r;
}
template <int N>
__device__ __forceinline__ DN<N> operator/(const DN<N> &a, const DN<N> &b) {
    DN<N> r;
    r.v = a.v / b.v;
#pragma unroll
    for (int k = 0; k < N; ++k) r.g[k] = (a.g[k] - r.v * b.g[k]) / b.v;
    return r;
}

// ---- the power as a function of its raw sums -------------------------------
template <class X>
__device__ __forceinline__ X gen_power(const X &W, const X &A1, const X &A2, const X &C,
                                       const X &S, const X &C2, const X &S2, const X &YC,
                                       const X &YS) {
    const X iw = 1.0 / W;
    const X cb = C * iw, sb = S * iw, yb = A1 * iw;
    const X cc = 0.5 + 0.5 * (C2 * iw) - cb * cb;
    const X ss = 0.5 + (-0.5) * (C2 * iw) - sb * sb;
    const X cs = 0.5 * (S2 * iw) - cb * sb;
    const X yc = YC * iw - yb * cb;
    const X ys = YS * iw - yb * sb;
    const X yy = A2 * iw - yb * yb;
    const X num = ss * yc * yc - 2.0 * (cs * yc * ys) + cc * ys * ys;
    const X den = (cc * ss - cs * cs) * yy;
    return num / den;
}

template <class X>
__device__ __forceinline__ X std_power(double N, const X &C2, const X &S2, const X &XC,
                                       const X &XS) {
    const X cc = 0.5 * (N + C2);
    const X ss = 0.5 * (N + (-1.0) * C2);
    const X cs = 0.5 * S2;
    const X num = ss * XC * XC - 2.0 * (cs * XC * XS) + cc * XS * XS;
    const X den = 2.0 * (cc * ss - cs * cs);
    return num / den;
}

template <>
__device__ __forceinline__ J2 gen_power<J2>(const J2 &W, const J2 &A1, const J2 &A2, const J2 &C,
                                            const J2 &S, const J2 &C2, const J2 &S2, const J2 &YC,
                                            const J2 &YS) {
    // W, A1, A2 do not depend on omega: divide by the scalar W directly.
    const double iw = 1.0 / W.v, yb = A1.v * iw;
    const J2 cb = iw * C, sb = iw * S;
    const J2 cc = 0.5 + (0.5 * iw) * C2 - cb * cb;
    const J2 ss = 0.5 + (-0.5 * iw) * C2 - sb * sb;
    const J2 cs = (0.5 * iw) * S2 - cb * sb;
    const J2 yc = iw * YC - yb * cb;
    const J2 ys = iw * YS - yb * sb;
    const double yy = A2.v * iw - yb * yb;
    const J2 num = ss * yc * yc - 2.0 * (cs * yc * ys) + cc * ys * ys;
    const J2 den = yy * (cc * ss - cs * cs);
    return num / den;
}

// ---- wave reductions --------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

struct PointArgs {
    const uint64_t *offsets;
    const void *t, *y, *dy;
    double *omega, *power, *dpower, *d2power;
    const double *grad_scale;
    void *grad_y, *grad_dy;
    double max_step, tol;
    uint32_t K;
    int32_t max_iter;
    int32_t *iterations;
    int32_t normalize;
    int32_t refine;
};

// Per-object constants of the pre-passes.
struct ObjStats {
    uint64_t lo, hi;
    double tref;   // middle of the time span
    double ybar;   // weighted (generalized) or plain (standard) mean of y
    double W;      // sum u (generalized) or N (standard)
    double A1;     // sum u (y - ybar): zero up to rounding
    double A2;     // sum u (y - ybar)^2
};

// Sums of one evaluation at omega, every lane holding the totals.
//   generalized: s[0..5]  = u c, u s, u c2, u s2, u yv c, u yv s
//   standard:    s[0..3]  = c2, s2, y c, y s
//   s[6..11] / s[4..7]: the same weighted by t, then by t^2.
template <bool GEN>
struct Sums {
    static constexpr int B = GEN ? 6 : 4;
    double s[3 * B];
};

template <typename R, bool GEN>
__device__ void eval_sums(const PointArgs &a, const ObjStats &st, double omega, int lane,
                          Sums<GEN> &out) {
    constexpr int B = Sums<GEN>::B;
    double acc[3 * B];
#pragma unroll
    for (int k = 0; k < 3 * B; ++k) acc[k] = 0.0;
    const R *t = static_cast<const R *>(a.t);
    const R *y = static_cast<const R *>(a.y);
    const R *dy = static_cast<const R *>(a.dy);
    for (uint64_t j = st.lo + lane; j < st.hi; j += 64) {
        const double tt = (double)t[j] - st.tref;
        double s, c;
        sincos(omega * tt, &s, &c);
        const double c2 = (c - s) * (c + s), s2 = 2.0 * s * c;
        double v[B];
        if constexpr (GEN) {
            const double d = (double)dy[j];
            const double u = 1.0 / (d * d);
            const double b = u * ((double)y[j] - st.ybar);
            v[0] = u * c;
            v[1] = u * s;
            v[2] = u * c2;
            v[3] = u * s2;
            v[4] = b * c;
            v[5] = b * s;
        } else {
            const double b = (double)y[j];
            v[0] = c2;
            v[1] = s2;
            v[2] = b * c;
            v[3] = b * s;
        }
        const double t2 = tt * tt;
#pragma unroll
        for (int k = 0; k < B; ++k) {
            acc[k] += v[k];
            acc[B + k] = fma(tt, v[k], acc[B + k]);
            acc[2 * B + k] = fma(t2, v[k], acc[2 * B + k]);
        }
    }
#pragma unroll
    for (int k = 0; k < 3 * B; ++k) out.s[k] = wave_sum(acc[k]);
}

// Jets of the raw sums in omega: for F = sum a cos(m w t), F' = -m sum a t sin,
// F'' = -m^2 sum a t^2 cos; for G = sum a sin(m w t), G' = m sum a t cos,
// G'' = -m^2 sum a t^2 sin.
template <bool GEN>
__device__ __forceinline__ J2 jet_cos(const Sums<GEN> &q, int ic, int is, double m) {
    constexpr int B = Sums<GEN>::B;
    return {q.s[ic], -m * q.s[B + is], -m * m * q.s[2 * B + ic]};
}
template <bool GEN>
__device__ __forceinline__ J2 jet_sin(const Sums<GEN> &q, int is, int ic, double m) {
    constexpr int B = Sums<GEN>::B;
    return {q.s[is], m * q.s[B + ic], -m * m * q.s[2 * B + is]};
}

template <bool GEN>
__device__ __forceinline__ J2 power_jet(const ObjStats &st, const Sums<GEN> &q, double scale) {
    if constexpr (GEN) {
        const J2 C = jet_cos(q, 0, 1, 1.0), S = jet_sin(q, 1, 0, 1.0);
        const J2 C2 = jet_cos(q, 2, 3, 2.0), S2 = jet_sin(q, 3, 2, 2.0);
        const J2 YC = jet_cos(q, 4, 5, 1.0), YS = jet_sin(q, 5, 4, 1.0);
        return gen_power<J2>(jc(st.W), jc(st.A1), jc(st.A2), C, S, C2, S2, YC, YS);
    } else {
        const J2 C2 = jet_cos(q, 0, 1, 2.0), S2 = jet_sin(q, 1, 0, 2.0);
        const J2 XC = jet_cos(q, 2, 3, 1.0), XS = jet_sin(q, 3, 2, 1.0);
        return scale * std_power<J2>(st.W, C2, S2, XC, XS);
    }
}

template <typename R, bool GEN>
__device__ ObjStats object_stats(const PointArgs &a, uint32_t k, int lane) {
    ObjStats st;
    st.lo = a.offsets[k];
    st.hi = a.offsets[k + 1];
    const R *t = static_cast<const R *>(a.t);
    const R *y = static_cast<const R *>(a.y);
    const R *dy = static_cast<const R *>(a.dy);
    double tmin = INFINITY, tmax = -INFINITY, w = 0.0, wy = 0.0;
    for (uint64_t j = st.lo + lane; j < st.hi; j += 64) {
        const double tj = (double)t[j];
        tmin = fmin(tmin, tj);
        tmax = fmax(tmax, tj);
        double u = 1.0;
        if constexpr (GEN) {
            const double d = (double)dy[j];
            u = 1.0 / (d * d);
        }
        w += u;
        wy += u * (double)y[j];
    }
    tmin = wave_min(tmin);
    tmax = wave_max(tmax);
    st.W = wave_sum(w);
    st.tref = 0.5 * (tmin + tmax);
    st.ybar = wave_sum(wy) / st.W;
    double a1 = 0.0, a2 = 0.0;
    for (uint64_t j = st.lo + lane; j < st.hi; j += 64) {
        double u = 1.0;
        if constexpr (GEN) {
            const double d = (double)dy[j];
            u = 1.0 / (d * d);
        }
        const double yv = (double)y[j] - st.ybar;
        a1 += u * yv;
        a2 += u * yv * yv;
    }
    st.A1 = wave_sum(a1);
    st.A2 = wave_sum(a2);
    return st;
}

// Per-observation gradients at omega, scaled by gs.
template <typename R, bool GEN>
__device__ void write_grads(const PointArgs &a, const ObjStats &st, const Sums<GEN> &q,
                            double omega, double scale, double P, double gs, int lane) {
    const R *t = static_cast<const R *>(a.t);
    const R *y = static_cast<const R *>(a.y);
    const R *dy = static_cast<const R *>(a.dy);
    R *gy = static_cast<R *>(a.grad_y);
    R *gd = static_cast<R *>(a.grad_dy);
    if constexpr (GEN) {
        // partials with respect to (W, A1, A2, C, S, C2, S2, YC, YS)
        const DN<9> P9 = gen_power<DN<9>>(dseed<9>(st.W, 0), dseed<9>(st.A1, 1), dseed<9>(st.A2, 2),
                                          dseed<9>(q.s[0], 3), dseed<9>(q.s[1], 4),
                                          dseed<9>(q.s[2], 5), dseed<9>(q.s[3], 6),
                                          dseed<9>(q.s[4], 7), dseed<9>(q.s[5], 8));
        double g[9];
#pragma unroll
        for (int i = 0; i < 9; ++i) g[i] = gs * P9.g[i];
        for (uint64_t j = st.lo + lane; j < st.hi; j += 64) {
            const double tt = (double)t[j] - st.tref;
            double s, c;
            sincos(omega * tt, &s, &c);
            const double c2 = (c - s) * (c + s), s2 = 2.0 * s * c;
            const double d = (double)dy[j];
            const double u = 1.0 / (d * d);
            const double yv = (double)y[j] - st.ybar;
            if (gy) gy[j] = (R)(u * (g[1] + 2.0 * g[2] * yv + g[7] * c + g[8] * s));
            if (gd) {
                const double du = g[0] + g[1] * yv + g[2] * yv * yv + g[3] * c + g[4] * s +
                                  g[5] * c2 + g[6] * s2 + g[7] * yv * c + g[8] * yv * s;
                gd[j] = (R)(du * (-2.0 * u / d));
            }
        }
    } else {
        // partials with respect to (XC, XS); C2, S2 do not depend on y
        const DN<2> XC = dseed<2>(q.s[2], 0), XS = dseed<2>(q.s[3], 1);
        const DN<2> C2 = dconst<2>(q.s[0]), S2 = dconst<2>(q.s[1]);
        const DN<2> P2 = std_power<DN<2>>(st.W, C2, S2, XC, XS);
        const double gxc = gs * scale * P2.g[0], gxs = gs * scale * P2.g[1];
        // normalisation 2/A2 with A2 = sum (y - ybar)^2: d(scale)/dy_j = -scale 2 (y_j - ybar)/A2
        const double gn = a.normalize ? gs * P * 2.0 / st.A2 : 0.0;
        if (!gy) return;
        for (uint64_t j = st.lo + lane; j < st.hi; j += 64) {
            const double tt = (double)t[j] - st.tref;
            double s, c;
            sincos(omega * tt, &s, &c);
            const double yv = (double)y[j] - st.ybar;
            gy[j] = (R)(gxc * c + gxs * s - gn * yv);
        }
        if (gd) {
            for (uint64_t j = st.lo + lane; j < st.hi; j += 64) gd[j] = (R)0;
        }
    }
}

template <typename R, bool GEN>
__global__ __launch_bounds__(256) void k_point(PointArgs a) {
    const int lane = threadIdx.x & 63;
    const uint32_t k = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (k >= a.K) return;  // whole wavefronts exit together
    const ObjStats st = object_stats<R, GEN>(a, k, lane);
    const double scale = (!GEN && a.normalize) ? 2.0 / st.A2 : 1.0;
    double omega = a.omega[k];
    Sums<GEN> q;
    int32_t iters = 0;
    if (a.refine) {
        bool converged = false;
        for (int it = 0; it < a.max_iter; ++it) {
            eval_sums<R, GEN>(a, st, omega, lane, q);
            const J2 P = power_jet<GEN>(st, q, scale);
            double step = -P.d / P.dd;
            // every lane computes the same step from the same totals
            if (!(P.dd < 0.0) || !isfinite(step) || fabs(step) > a.max_step)
                step = (P.d == 0.0 || !isfinite(P.d)) ? 0.0 : copysign(a.max_step, P.d);
            omega += step;
            iters = it + 1;
            if (fabs(step) < a.tol) {
                converged = true;
                break;
            }
        }
        if (!converged) iters = -iters;
    }
    eval_sums<R, GEN>(a, st, omega, lane, q);
    const J2 P = power_jet<GEN>(st, q, scale);
    if (lane == 0) {
        if (a.refine) a.omega[k] = omega;
        if (a.iterations) a.iterations[k] = iters;
        if (a.power) a.power[k] = P.v;
        if (a.dpower) a.dpower[k] = P.d;
        if (a.d2power) a.d2power[k] = P.dd;
    }
    if (a.grad_y || a.grad_dy) {
        const double gs = a.grad_scale ? a.grad_scale[k] : 1.0;
        write_grads<R, GEN>(a, st, q, omega, scale, P.v, gs, lane);
    }
}

}  // namespace

#define LS_HIP(call)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess)                                                                \
            return set_error(LSAMD_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

static int point_launch(const lsamd_point_args *a, void *stream_v, bool refine, const char *fn) {
    const std::string name(fn);
    if (!a) return set_error(LSAMD_ERR_INVALID, name + ": null args");
    if (a->variant != LSAMD_STANDARD && a->variant != LSAMD_GENERALIZED)
        return set_error(LSAMD_ERR_INVALID, name + ": bad variant");
    if (a->dtype != LSAMD_F32 && a->dtype != LSAMD_F64)
        return set_error(LSAMD_ERR_INVALID, name + ": bad dtype");
    if (a->num_objects == 0 || a->num_obs == 0)
        return set_error(LSAMD_ERR_INVALID, name + ": empty input");
    if (!a->offsets || !a->t || !a->y || !a->omega)
        return set_error(LSAMD_ERR_INVALID, name + ": null device pointer");
    const bool gen = a->variant == LSAMD_GENERALIZED;
    if (gen && !a->dy) return set_error(LSAMD_ERR_INVALID, name + ": generalized variant needs dy");
    if (refine) {
        if (!(a->max_step > 0.0) || !(a->tol > 0.0) || a->max_iter < 1)
            return set_error(LSAMD_ERR_INVALID, name + ": need max_step > 0, tol > 0, max_iter >= 1");
    } else if (!a->power && !a->dpower && !a->d2power && !a->grad_y && !a->grad_dy) {
        return set_error(LSAMD_ERR_INVALID, name + ": nothing to compute");
    }
    PointArgs p{};
    p.offsets = a->offsets;
    p.t = a->t;
    p.y = a->y;
    p.dy = a->dy;
    p.omega = a->omega;
    p.power = a->power;
    p.dpower = a->dpower;
    p.d2power = a->d2power;
    p.grad_scale = a->grad_scale;
    p.grad_y = a->grad_y;
    p.grad_dy = a->grad_dy;
    p.max_step = a->max_step;
    p.tol = a->tol;
    p.K = a->num_objects;
    p.max_iter = refine ? a->max_iter : 0;
    p.iterations = refine ? a->iterations : nullptr;
    p.normalize = a->normalize ? 1 : 0;
    p.refine = refine ? 1 : 0;
    hipStream_t st = static_cast<hipStream_t>(stream_v);
    const int waves = 4;
    const dim3 block(64 * waves), grid((a->num_objects + waves - 1) / waves);
    void (*kern)(PointArgs) =
        a->dtype == LSAMD_F64 ? (gen ? k_point<double, true> : k_point<double, false>)
                              : (gen ? k_point<float, true> : k_point<float, false>);
    hipLaunchKernelGGL(kern, grid, block, 0, st, p);
    LS_HIP(hipGetLastError());
    return LSAMD_OK;
}

}  // namespace lsamd

extern "C" {

int lsamd_power_grad(const lsamd_point_args *a, void *stream) {
    return lsamd::point_launch(a, stream, false, "lsamd_power_grad");
}

int lsamd_refine_peaks(const lsamd_point_args *a, void *stream) {
    return lsamd::point_launch(a, stream, true, "lsamd_refine_peaks");
}

}  // extern "C"
