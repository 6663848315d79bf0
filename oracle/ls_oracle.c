/*
 * CPU oracle for the Lomb-Scargle hot path: a plain-C restatement of the
 * reference's CPU path (source/main.cu, modes 4/5), in float and double.
 *
 * TEST INFRASTRUCTURE ONLY.  The GPU library never links, loads or calls
 * this file; only tests/, the smoke check and the benchmark's CPU-baseline
 * leg use it, to check GPU results against.
 *
 * How it is pinned (each check lives in tests/):
 *   - bit-for-bit against the reference's own CPU functions compiled from
 *     their sources into oracle/_ref/ (when the reference tree is present),
 *     and against golden vectors produced that way (tests/golden/);
 *   - against the reference's recorded known answers (README transcripts and
 *     the Python example's recorded output);
 *   - against closed forms: the standard power equals 1/2 y'A(A'A)^-1 A'y
 *     and the generalized power equals 1 - chi2_fit/chi2_0 of a weighted
 *     a*cos + b*sin + c fit (brute-force least squares in numpy), plus
 *     invariances (time shift, permutation, scaling) and scipy's periodogram.
 *
 * Build: cc -O2 -fPIC -shared -ffp-contract=off -fopenmp (see oracle/Makefile).
 * -ffp-contract=off keeps every multiply and add separately rounded, as the
 * reference's x86-64 host build does.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* Runs of equal object ids are objects; a repeated non-contiguous id starts
 * a new object.  Returns the number of objects K and, when the arrays are
 * non-NULL, each object's inclusive index range and id.
 * Restates computeObjectRanges, source/main.cu:468-510. */
unsigned lso_object_ranges(const unsigned *objId, unsigned n, unsigned *idxMin,
                           unsigned *idxMax, unsigned *ids)
{
    unsigned k = 0, last;
    if (n == 0)
        return 0;
    last = objId[0];
    if (idxMin) idxMin[0] = 0;
    for (unsigned i = 1; i < n; i++) {
        if (objId[i] != last) {
            if (ids) ids[k] = last;
            if (idxMax) idxMax[k] = i - 1;
            if (idxMin) idxMin[k + 1] = i;
            k++;
            last = objId[i];
        }
    }
    if (ids) ids[k] = objId[n - 1];
    if (idxMax) idxMax[k] = n - 1;
    return k + 1;
}

#define CAT_(a, b) a##_##b
#define CAT(a, b) CAT_(a, b)
#define FN(name) CAT(name, SFX)

/* double instantiation */
#define REAL double
#define SFX d
#define SINR sin
#define COSR cos
#define ATAN2R atan2
#define SQRTR sqrt
#include "ls_oracle_body.h"
#undef REAL
#undef SFX
#undef SINR
#undef COSR
#undef ATAN2R
#undef SQRTR

/* float instantiation: the reference's C++ resolves sin/cos/atan2/sqrt of a
 * float argument to the float overloads, i.e. glibc's sinf/cosf/atan2f.
 * lso_set_float_libm_cr(1) swaps those three for correctly rounded versions
 * (evaluated in double, rounded once) so tests can measure how much the
 * reference's float algorithm depends on its libm; default 0 = glibc. */
static int g_float_cr = 0;
void lso_set_float_libm_cr(int on) { g_float_cr = on != 0; }
static float sinf_sel(float x) { return g_float_cr ? (float)sin((double)x) : sinf(x); }
static float cosf_sel(float x) { return g_float_cr ? (float)cos((double)x) : cosf(x); }
static float atan2f_sel(float y, float x) { return g_float_cr ? (float)atan2((double)y, (double)x) : atan2f(y, x); }
#define REAL float
#define SFX f
#define SINR sinf_sel
#define COSR cosf_sel
#define ATAN2R atan2f_sel
#define SQRTR sqrtf
#include "ls_oracle_body.h"
#undef REAL
#undef SFX
#undef SINR
#undef COSR
#undef ATAN2R
#undef SQRTR
