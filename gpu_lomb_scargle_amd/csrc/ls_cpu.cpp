// CPU modes 4 (batch) and 5 (single object) of the drop-in library: the
// reference's CPU path with the same arithmetic, so that a caller asking
// for "CPU" gets the reference's numbers.  Used only when the caller selects
// a CPU mode; the GPU modes never fall back to it.
//
// Follows lombscarglecpuinnerloop (source/main.cu:1566-1599),
// lombscarglecpuinnerloopAstroPy (main.cu:1701-1777), updateYerrorfactor
// (main.cu:1631-1662), computeStandardDevSquared / normalizepgram*
// (main.cu:1295-1352) and the drivers (main.cu:1358-1543).  float and double
// are template instances; where the reference's C++ promotes a float to
// double (a double literal, a double-only libm call) the promotion is
// written out.  Built with -ffp-contract=off so products and sums round
// separately, like the reference's x86-64 host build.

#include <cmath>
#include <cstdint>
#include <omp.h>
#include <vector>

#include "ls_internal.h"

namespace lsamd {
namespace {

inline float l_sin(float x) { return sinf(x); }
inline double l_sin(double x) { return sin(x); }
inline float l_cos(float x) { return cosf(x); }
inline double l_cos(double x) { return cos(x); }
inline float l_atan2(float y, float x) { return atan2f(y, x); }
inline double l_atan2(double y, double x) { return atan2(y, x); }

template <typename R>
R power_std(R freq, const R *x, const R *y, uint64_t n) {
    R xc = 0, xs = 0, cc = 0, ss = 0, cs = 0;
    for (uint64_t j = 0; j < n; ++j) {
        const R c = l_cos(freq * x[j]);
        const R s = l_sin(freq * x[j]);
        xc += y[j] * c;
        xs += y[j] * s;
        cc += c * c;
        ss += s * s;
        cs += c * s;
    }
    const R tau = (R)(atan2(2.0 * (double)cs, (double)(cc - ss)) / (2.0 * (double)freq));
    const R c_tau = l_cos(freq * tau);
    const R s_tau = l_sin(freq * tau);
    const R c_tau2 = c_tau * c_tau, s_tau2 = s_tau * s_tau;
    const R cs_tau = (R)(2.0 * (double)c_tau * (double)s_tau);
    const R a = c_tau * xc + s_tau * xs;
    const R b = c_tau * xs - s_tau * xc;
    const R q = (a * a) / (c_tau2 * cc + cs_tau * cs + s_tau2 * ss) +
                (b * b) / (c_tau2 * ss - cs_tau * cs + s_tau2 * cc);
    return (R)(0.5 * (double)q);
}

template <typename R>
R power_gen(R freq, const R *x, const R *y, const R *dy, uint64_t n) {
    R S = 0, C = 0, S2 = 0, C2 = 0, wsum = 0;
    for (uint64_t j = 0; j < n; ++j) {
        R w = (R)(1.0 / (double)dy[j]);
        w *= w;
        wsum += w;
        const R ot = freq * x[j];
        const R so = l_sin(ot), co = l_cos(ot);
        S += w * so;
        C += w * co;
        S2 = (R)((double)S2 + 2.0 * (double)w * (double)so * (double)co);
        C2 = (R)((double)C2 + ((double)w - 2.0 * (double)w * (double)so * (double)so));
    }
    S2 /= wsum;
    C2 /= wsum;
    S /= wsum;
    C /= wsum;
    S2 = (R)((double)S2 - 2.0 * (double)S * (double)C);
    C2 -= (C * C - S * S);
    const R tau = (R)(0.5 * (double)l_atan2(S2, C2) / (double)freq);
    R Y = 0, YY = 0, St = 0, Ct = 0, YC = 0, YS = 0, CC = 0, SS = 0;
    for (uint64_t j = 0; j < n; ++j) {
        R w = (R)(1.0 / (double)dy[j]);
        w *= w;
        const R ot = freq * (x[j] - tau);
        const R so = l_sin(ot), co = l_cos(ot);
        Y += w * y[j];
        YY += w * y[j] * y[j];
        Ct += w * co;
        St += w * so;
        YC += w * y[j] * co;
        YS += w * y[j] * so;
        CC += w * co * co;
        SS += w * so * so;
    }
    Y /= wsum;
    YY /= wsum;
    Ct /= wsum;
    St /= wsum;
    YC /= wsum;
    YS /= wsum;
    CC /= wsum;
    SS /= wsum;
    YC -= Y * Ct;
    YS -= Y * St;
    CC -= Ct * Ct;
    SS -= St * St;
    YY -= Y * Y;
    return (YC * YC / CC + YS * YS / SS) / YY;
}

template <typename R>
void normalize_obj(R *pg, const R *y, uint64_t n, uint64_t nf) {
    double sum = 0, sum2 = 0;
    for (uint64_t k = 0; k < n; ++k) sum += y[k];
    const double mean = sum / ((double)n * 1.0);
    for (uint64_t k = 0; k < n; ++k) sum2 += ((double)y[k] - mean) * ((double)y[k] - mean);
    const double sigma = sqrt(sum2 / ((double)n * 1.0));
    const double sd2 = sigma * sigma;
    for (uint64_t j = 0; j < nf; ++j) pg[j] = (R)((double)pg[j] * (2.0 / ((double)n * sd2)));
}

template <typename R>
void sweep(int gen, R fmin, R fstep, uint32_t nf, const R *x, const R *y, const R *dy, uint64_t n,
           R *pg, int nthreads) {
#pragma omp parallel for schedule(static) num_threads(nthreads) if (nthreads > 1)
    for (long long i = 0; i < (long long)nf; ++i) {
        const R f = fmin + (fstep * (R)(uint32_t)i);
        pg[i] = gen ? power_gen(f, x, y, dy, n) : power_std(f, x, y, n);
    }
}

int resolve_threads(int nthreads) {
    if (nthreads > 0) return nthreads;
    return 0;  // 0: OpenMP default
}

template <typename R>
int batch_t(int variant, int normalize, const uint32_t *id, const R *t, R *y, const R *dy,
            uint64_t n, double fminD, double fmaxD, uint32_t nf, R *pg, int nthreads) {
    const uint32_t K = count_runs(id, n);
    std::vector<uint64_t> off((size_t)K + 1);
    run_offsets(id, n, off.data(), nullptr);
    const R fmin = (R)fminD, fmax = (R)fmaxD;
    const R fstep = (R)((double)(fmax - fmin) / ((double)nf * 1.0));
    const bool gen = variant == LSAMD_GENERALIZED;
    const int nt = resolve_threads(nthreads);
#pragma omp parallel for schedule(dynamic) num_threads(nt > 0 ? nt : omp_get_max_threads())
    for (long long k = 0; k < (long long)K; ++k) {
        const uint64_t a = off[k], m = off[k + 1] - off[k];
        R *p = pg + (uint64_t)k * nf;
        if (gen) precenter_host(y + a, dy + a, m);
        sweep<R>(gen, fmin, fstep, nf, t + a, y + a, gen ? dy + a : nullptr, m, p, 1);
        if (normalize && !gen) normalize_obj(p, y + a, m, nf);
    }
    return LSAMD_OK;
}

template <typename R>
int single_t(int variant, int normalize, const R *t, R *y, const R *dy, uint64_t n, double fminD,
             double fmaxD, uint32_t nf, R *pg, int nthreads) {
    const R fmin = (R)fminD, fmax = (R)fmaxD;
    const R fstep = (R)((double)(fmax - fmin) / ((double)nf * 1.0));
    const bool gen = variant == LSAMD_GENERALIZED;
    if (gen) precenter_host(y, dy, n);
    const int nt = resolve_threads(nthreads);
    sweep<R>(gen, fmin, fstep, nf, t, y, dy, n, pg, nt > 0 ? nt : omp_get_max_threads());
    if (normalize && !gen) normalize_obj(pg, y, n, nf);
    return LSAMD_OK;
}

}  // namespace

int cpu_batch(int variant, int dtype, int normalize, const uint32_t *id, const void *t, void *y,
              const void *dy, uint64_t n, double fmin, double fmax, uint32_t nf, void *pgram,
              int nthreads) {
    if (n == 0 || nf == 0 || !id || !t || !y || !pgram ||
        (variant == LSAMD_GENERALIZED && !dy))
        return set_error(LSAMD_ERR_INVALID, "lsamd_cpu_batch: empty input or null array");
    if (dtype == LSAMD_F64)
        return batch_t<double>(variant, normalize, id, (const double *)t, (double *)y,
                               (const double *)dy, n, fmin, fmax, nf, (double *)pgram, nthreads);
    return batch_t<float>(variant, normalize, id, (const float *)t, (float *)y, (const float *)dy, n,
                          fmin, fmax, nf, (float *)pgram, nthreads);
}

int cpu_single(int variant, int dtype, int normalize, const void *t, void *y, const void *dy,
               uint64_t n, double fmin, double fmax, uint32_t nf, void *pgram, int nthreads) {
    if (n == 0 || nf == 0 || !t || !y || !pgram || (variant == LSAMD_GENERALIZED && !dy))
        return set_error(LSAMD_ERR_INVALID, "lsamd_cpu_single: empty input or null array");
    if (dtype == LSAMD_F64)
        return single_t<double>(variant, normalize, (const double *)t, (double *)y,
                                (const double *)dy, n, fmin, fmax, nf, (double *)pgram, nthreads);
    return single_t<float>(variant, normalize, (const float *)t, (float *)y, (const float *)dy, n,
                           fmin, fmax, nf, (float *)pgram, nthreads);
}

}  // namespace lsamd
