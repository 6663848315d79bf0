/*
 * Type-generic body of the CPU oracle.  Included twice by ls_oracle.c, once
 * with REAL=double and once with REAL=float.  Every function restates one
 * function of the reference CPU path (source/main.cu), keeping the
 * reference's evaluation order and its float->double promotions, so that the
 * float build reproduces the reference float build and the double build the
 * double one.  Promotions that the reference's C++ performs implicitly (a
 * double literal in a float expression, a float-overloaded libm call) are
 * written out explicitly here.
 *
 * Test infrastructure only: nothing outside tests/, the smoke check and the
 * CPU-baseline leg of the benchmark may call it.
 */

/* Standard (un-weighted) L-S power at one angular frequency.
 * Restates lombscarglecpuinnerloop, source/main.cu:1566-1599. */
REAL FN(lso_std_power)(REAL freq, const REAL *x, const REAL *y, unsigned n)
{
    REAL c, s, xc = 0, xs = 0, cc = 0, ss = 0, cs = 0;
    REAL tau, c_tau, s_tau, c_tau2, s_tau2, cs_tau;
    for (unsigned j = 0; j < n; j++) {
        c = COSR(freq * x[j]);
        s = SINR(freq * x[j]);
        xc += y[j] * c;
        xs += y[j] * s;
        cc += c * c;
        ss += s * s;
        cs += c * s;
    }
    /* main.cu:1588: atan2(2.0*cs, cc-ss) is a double atan2 in both builds */
    tau = (REAL)(atan2(2.0 * (double)cs, (double)(cc - ss)) / (2.0 * (double)freq));
    c_tau = COSR(freq * tau);
    s_tau = SINR(freq * tau);
    c_tau2 = c_tau * c_tau;
    s_tau2 = s_tau * s_tau;
    cs_tau = (REAL)(2.0 * (double)c_tau * (double)s_tau);
    {
        REAL a = c_tau * xc + s_tau * xs;
        REAL b = c_tau * xs - s_tau * xc;
        REAL q = (a * a) / (c_tau2 * cc + cs_tau * cs + s_tau2 * ss)
               + (b * b) / (c_tau2 * ss - cs_tau * cs + s_tau2 * cc);
        return (REAL)(0.5 * (double)q);
    }
}

/* Generalized (floating-mean, 1/dy^2-weighted) power at one frequency.
 * Restates lombscarglecpuinnerloopAstroPy, source/main.cu:1701-1777
 * (two passes: tau, then the tau-shifted sums). */
REAL FN(lso_gen_power)(REAL freq, const REAL *x, const REAL *y, const REAL *dy,
                       unsigned n)
{
    REAL w, omega_t, so, co, S = 0, C = 0, S2 = 0, C2 = 0, tau, wsum = 0;
    REAL Y = 0, YY = 0, Stau = 0, Ctau = 0, YCtau = 0, YStau = 0, CCtau = 0, SStau = 0;
    for (unsigned j = 0; j < n; j++) {
        w = (REAL)(1.0 / (double)dy[j]);
        w *= w;
        wsum += w;
        omega_t = freq * x[j];
        so = SINR(omega_t);
        co = COSR(omega_t);
        S += w * so;
        C += w * co;
        /* main.cu:1726-1727: 2.0*w*s*c is evaluated in double */
        S2 = (REAL)((double)S2 + 2.0 * (double)w * (double)so * (double)co);
        C2 = (REAL)((double)C2 + ((double)w - 2.0 * (double)w * (double)so * (double)so));
    }
    S2 /= wsum;
    C2 /= wsum;
    S /= wsum;
    C /= wsum;
    S2 = (REAL)((double)S2 - 2.0 * (double)S * (double)C);
    C2 -= (C * C - S * S);
    /* main.cu:1734: atan2(S2, C2) takes two DTYPE arguments, so the float
     * build resolves to the float overload */
    tau = (REAL)(0.5 * (double)ATAN2R(S2, C2) / (double)freq);
    for (unsigned j = 0; j < n; j++) {
        w = (REAL)(1.0 / (double)dy[j]);
        w *= w;
        omega_t = freq * (x[j] - tau);
        so = SINR(omega_t);
        co = COSR(omega_t);
        Y += w * y[j];
        YY += w * y[j] * y[j];
        Ctau += w * co;
        Stau += w * so;
        YCtau += w * y[j] * co;
        YStau += w * y[j] * so;
        CCtau += w * co * co;
        SStau += w * so * so;
    }
    Y /= wsum;
    YY /= wsum;
    Ctau /= wsum;
    Stau /= wsum;
    YCtau /= wsum;
    YStau /= wsum;
    CCtau /= wsum;
    SStau /= wsum;
    YCtau -= Y * Ctau;
    YStau -= Y * Stau;
    CCtau -= Ctau * Ctau;
    SStau -= Stau * Stau;
    YY -= Y * Y;
    return (YCtau * YCtau / CCtau + YStau * YStau / SStau) / YY;
}

/* In-place pre-centring of the magnitudes before the generalized sweep,
 * with the reference's weights w = 1/sqrt(dy) (not dy^-2).
 * Restates updateYerrorfactor, source/main.cu:1631-1662; the reference's
 * OpenMP reductions are summed here in index order. */
void FN(lso_precenter)(REAL *y, const REAL *dy, unsigned n)
{
    REAL sumw = 0, dotwy = 0;
    for (unsigned i = 0; i < n; i++)
        sumw += (REAL)(1.0 / (double)SQRTR(dy[i]));
    for (unsigned i = 0; i < n; i++)
        dotwy += (REAL)(1.0 / (double)SQRTR(dy[i])) * y[i];
    for (unsigned i = 0; i < n; i++)
        y[i] = y[i] - dotwy / sumw;
}

/* Population variance (two-pass, double).
 * Restates computeStandardDevSquared, source/main.cu:1295-1317. */
double FN(lso_stddev_sq)(const REAL *y, uint64_t n)
{
    double sum = 0, sum2 = 0, mean, sigma;
    for (uint64_t k = 0; k < n; k++)
        sum += y[k];
    mean = sum / ((double)n * 1.0);
    for (uint64_t k = 0; k < n; k++)
        sum2 += ((double)y[k] - mean) * ((double)y[k] - mean);
    sigma = sqrt(sum2 / ((double)n * 1.0));
    return sigma * sigma;
}

/* pgram *= 2/(N sigma^2), standard variant in the Python builds.
 * Restates normalizepgramsingleobject, source/main.cu:1341-1352. */
void FN(lso_normalize)(REAL *pgram, const REAL *y, uint64_t n, uint64_t nf)
{
    double sd2 = FN(lso_stddev_sq)(y, n);
    for (uint64_t j = 0; j < nf; j++)
        pgram[j] = (REAL)((double)pgram[j] * (2.0 / ((double)n * sd2)));
}

/* First index of the strict maximum, starting from power 0, and its period.
 * Restates computePeriod, source/main.cu:1545-1563. */
void FN(lso_compute_period)(const REAL *pgram, unsigned nf, REAL minFreq, REAL freqStep,
                            REAL *period, REAL *power, unsigned *index)
{
    REAL maxPower = 0;
    unsigned maxIdx = 0;
    for (unsigned i = 0; i < nf; i++) {
        if (maxPower < pgram[i]) {
            maxPower = pgram[i];
            maxIdx = i;
        }
    }
    *period = (REAL)((1.0 / (double)(minFreq + ((REAL)maxIdx * freqStep))) * 2.0 * M_PI);
    *power = maxPower;
    if (index)
        *index = maxIdx;
}

/* Frequency grid of the CPU path: step = (max-min)/Nf in DTYPE (max excluded),
 * omega_i = min + step*i in DTYPE.  source/main.cu:1371, 1611. */
REAL FN(lso_freq_step)(REAL minFreq, REAL maxFreq, unsigned nf)
{
    return (REAL)((double)(maxFreq - minFreq) / ((double)nf * 1.0));
}

static void FN(lso_sweep)(int error, REAL minFreq, REAL freqStep, unsigned nf,
                          const REAL *x, const REAL *y, const REAL *dy, unsigned n,
                          REAL *pgram, int nthreads)
{
    long long i;
    (void)nthreads;
#pragma omp parallel for schedule(static) num_threads(nthreads) if (nthreads > 1)
    for (i = 0; i < (long long)nf; i++) {
        REAL f = minFreq + (freqStep * (REAL)(unsigned)i);
        pgram[i] = error ? FN(lso_gen_power)(f, x, y, dy, n) : FN(lso_std_power)(f, x, y, n);
    }
}

/* Batch of objects = reference MODE 4 (CPU batch).
 * Restates lombscargleCPUBatch / lombscargleCPUBatchError,
 * source/main.cu:1401-1543: objects are contiguous runs of objId
 * (computeObjectRanges, main.cu:468-510); each object's y is pre-centred in
 * place for the generalized variant; the standard pgram is normalized when
 * `normalize` is set (NORMALIZEPGRAM=1 of the Python builds).  Objects are
 * spread over `nthreads` OpenMP threads (dynamic schedule, as in the
 * reference); the per-object arithmetic does not depend on the thread count.
 * periods/powers/indices (each K entries, may be NULL) follow computePeriod. */
void FN(lso_batch)(int error, int normalize, const unsigned *objId, const REAL *t, REAL *y,
                   const REAL *dy, unsigned n, double minFreqD, double maxFreqD, unsigned nf,
                   REAL *pgram, REAL *periods, REAL *powers, unsigned *indices, int nthreads)
{
    unsigned K = lso_object_ranges(objId, n, NULL, NULL, NULL);
    unsigned *lo = (unsigned *)malloc(sizeof(unsigned) * K);
    unsigned *hi = (unsigned *)malloc(sizeof(unsigned) * K);
    const REAL minFreq = (REAL)minFreqD, maxFreq = (REAL)maxFreqD;
    const REAL freqStep = FN(lso_freq_step)(minFreq, maxFreq, nf);
    long long k;
    lso_object_ranges(objId, n, lo, hi, NULL);
#pragma omp parallel for schedule(dynamic) num_threads(nthreads) if (nthreads > 1)
    for (k = 0; k < (long long)K; k++) {
        unsigned a = lo[k], m = hi[k] - lo[k] + 1;
        REAL *pg = pgram + (uint64_t)k * nf;
        REAL per, pw;
        unsigned ix;
        if (error)
            FN(lso_precenter)(y + a, dy + a, m);
        FN(lso_sweep)(error, minFreq, freqStep, nf, t + a, y + a, error ? dy + a : NULL, m, pg, 1);
        if (normalize && !error)
            FN(lso_normalize)(pg, y + a, m, nf);
        FN(lso_compute_period)(pg, nf, minFreq, freqStep, &per, &pw, &ix);
        if (periods) periods[k] = per;
        if (powers) powers[k] = pw;
        if (indices) indices[k] = ix;
    }
    free(lo);
    free(hi);
}

/* Whole input as one object = reference MODE 5 (CPU single object).
 * Restates lombscargleCPUOneObject / lombscargleCPUOneObjectError,
 * source/main.cu:1358-1398 (frequencies spread over threads, static). */
void FN(lso_single)(int error, int normalize, const REAL *t, REAL *y, const REAL *dy, unsigned n,
                    double minFreqD, double maxFreqD, unsigned nf, REAL *pgram, REAL *period,
                    REAL *power, unsigned *index, int nthreads)
{
    const REAL minFreq = (REAL)minFreqD, maxFreq = (REAL)maxFreqD;
    const REAL freqStep = FN(lso_freq_step)(minFreq, maxFreq, nf);
    if (error)
        FN(lso_precenter)(y, dy, n);
    FN(lso_sweep)(error, minFreq, freqStep, nf, t, y, dy, n, pgram, nthreads);
    if (normalize && !error)
        FN(lso_normalize)(pgram, y, n, nf);
    FN(lso_compute_period)(pgram, nf, minFreq, freqStep, period, power, index);
}

/* Period of frequency bin `idx` as the GPU host path computes it:
 * (1/(min + idx*step))*2*pi with step = (max-min)/Nf in double and
 * max-min in DTYPE.  source/main.cu:1136-1153. */
REAL FN(lso_period_of_index)(double minFreqD, double maxFreqD, unsigned nf, uint64_t idx)
{
    const REAL minFreq = (REAL)minFreqD, maxFreq = (REAL)maxFreqD;
    double step = (double)(maxFreq - minFreq) / ((double)nf * 1.0);
    return (REAL)((1.0 / ((double)minFreq + ((double)idx * step))) * 2.0 * M_PI);
}

/* Top three distinct bins per object: three scans, each the first strict
 * maximum from 0 excluding the bins already taken.
 * Restates the PRINTPERIODS==3 block, source/main.cu:1165-1219. */
void FN(lso_top3)(const REAL *pgram, unsigned K, unsigned nf, double minFreqD, double maxFreqD,
                  REAL *periods3, REAL *powers3, unsigned *idx3)
{
    for (uint64_t i = 0; i < K; i++) {
        const REAL *p = pgram + i * (uint64_t)nf;
        REAL mp[3] = {0, 0, 0};
        uint64_t mi[3] = {0, 0, 0};
        for (uint64_t j = 0; j < nf; j++)
            if (mp[0] < p[j]) { mp[0] = p[j]; mi[0] = j; }
        for (uint64_t j = 0; j < nf; j++)
            if (j != mi[0] && mp[1] < p[j]) { mp[1] = p[j]; mi[1] = j; }
        for (uint64_t j = 0; j < nf; j++)
            if (j != mi[0] && j != mi[1] && mp[2] < p[j]) { mp[2] = p[j]; mi[2] = j; }
        for (int r = 0; r < 3; r++) {
            periods3[i * 3 + r] = FN(lso_period_of_index)(minFreqD, maxFreqD, nf, mi[r]);
            powers3[i * 3 + r] = mp[r];
            idx3[i * 3 + r] = (unsigned)mi[r];
        }
    }
}
