// C entry points over the reference's own CPU functions.
//
// oracle/build_ref.sh compiles this file in ONE translation unit after the
// reference's CPU-path functions, which it reads by line range from
// source/main.cu under the reference tree (nothing of the reference is kept
// in this repository).  The resulting libraries land in oracle/_ref/ and are
// used only by tests/ (to pin the C oracle bit for bit) and, optionally, as
// the benchmark's CPU baseline.  DTYPE, ERROR, NORMALIZEPGRAM and NTHREADSCPU
// come from -D flags, as in the reference's own Python-library build
// (source/Makefile:47-91).

extern "C" {

// Reference MODE 4: lombscargleCPUBatch[Error] (main.cu:1401-1543).  The
// reference keeps its periods in a private buffer under -DPYTHON, so they are
// recomputed here with its computePeriod (main.cu:1545-1563).
void ref_batch(unsigned *objId, DTYPE *t, DTYPE *y, DTYPE *dy, unsigned n, double minFreq,
               double maxFreq, unsigned nf, DTYPE *pgram, DTYPE *periods, DTYPE *powers)
{
    DTYPE sumPeriods = 0;
#if ERROR == 0
    (void)dy;
    lombscargleCPUBatch(objId, t, y, &n, minFreq, maxFreq, nf, &sumPeriods, pgram, NULL, NULL);
#else
    lombscargleCPUBatchError(objId, t, y, dy, &n, minFreq, maxFreq, nf, &sumPeriods, pgram, NULL,
                             NULL);
#endif
    struct lookupObj *lk = NULL;
    unsigned K = 0;
    computeObjectRanges(objId, &n, &lk, &K);
    const DTYPE fmin = minFreq, fmax = maxFreq;
    const DTYPE step = (fmax - fmin) / (nf * 1.0);
    for (unsigned i = 0; i < K; i++)
        computePeriod(pgram + (uint64_t)i * nf, nf, fmin, step, &periods[i], &powers[i]);
    free(lk);
}

// Reference MODE 5: lombscargleCPUOneObject[Error] (main.cu:1358-1398).
void ref_single(DTYPE *t, DTYPE *y, DTYPE *dy, unsigned n, double minFreq, double maxFreq,
                unsigned nf, DTYPE *pgram, DTYPE *period, DTYPE *power)
{
#if ERROR == 0
    (void)dy;
    lombscargleCPUOneObject(t, y, &n, minFreq, maxFreq, nf, period, power, pgram);
#else
    lombscargleCPUOneObjectError(t, y, dy, &n, minFreq, maxFreq, nf, period, power, pgram);
#endif
}

// One power value: lombscarglecpuinnerloop (main.cu:1566) or
// lombscarglecpuinnerloopAstroPy (main.cu:1701).
DTYPE ref_power(DTYPE freq, DTYPE *x, DTYPE *y, DTYPE *dy, unsigned n)
{
    DTYPE out = 0;
#if ERROR == 0
    (void)dy;
    lombscarglecpuinnerloop(0, x, y, &out, &freq, n);
#else
    lombscarglecpuinnerloopAstroPy(0, x, y, dy, &out, &freq, n);
#endif
    return out;
}

// updateYerrorfactor (main.cu:1631-1662).
void ref_precenter(DTYPE *y, DTYPE *dy, unsigned n) { updateYerrorfactor(y, dy, n); }

// computeStandardDevSquared (main.cu:1295-1317).
double ref_stddev_sq(DTYPE *y, unsigned n) { return computeStandardDevSquared(y, n); }

// computeObjectRanges (main.cu:468-510).
unsigned ref_object_ranges(unsigned *objId, unsigned n, unsigned *lo, unsigned *hi, unsigned *ids)
{
    struct lookupObj *lk = NULL;
    unsigned K = 0;
    computeObjectRanges(objId, &n, &lk, &K);
    for (unsigned i = 0; i < K; i++) {
        lo[i] = lk[i].idxMin;
        hi[i] = lk[i].idxMax;
        ids[i] = lk[i].objId;
    }
    free(lk);
    return K;
}

int ref_nthreads(void) { return NTHREADSCPU; }
int ref_dtype_bytes(void) { return (int)sizeof(DTYPE); }

}  // extern "C"
