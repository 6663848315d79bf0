// The drop-in entry point LombScarglePy, compiled once per library variant
// (LSAMD_ABI_ERROR = 0/1, LSAMD_ABI_DTYPE = float/double) into
// libpyls{noerror,error}{float,double}.so, the names the reference's Python
// wrapper loads (source/Makefile:71-91, source/lombscarglegpu.py:279-287).
//
// Replaces LombScarglePy, source/main.cu:255-411: same signature, MODE
// dispatch (1 GPU batch, 2 GPU single object, 4 CPU batch, 5 CPU single
// object, anything else does nothing), caller-owned pgram fully
// overwritten, standard variant normalised by 2/(N sigma^2) as in the
// reference's Python builds (NORMALIZEPGRAM=1, source/Makefile:65),
// generalized variant pre-centring magY in place, progress lines on stdout
// and one appended record in ./gpu_stats.txt per call.  Unlike the
// reference, a failure is reported on stderr and through lsamd_last_error()
// and the call returns; LSAMD_EXIT_ON_ERROR=1 exits like the reference's
// gpuErrchk (main.cu:38-46).

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#ifdef LSAMD_ABI_DTYPE
#define LSAMD_DTYPE LSAMD_ABI_DTYPE  // the header's LombScarglePy prototype uses it
#endif
#include "ls_internal.h"

#ifndef LSAMD_ABI_ERROR
#error "LSAMD_ABI_ERROR must be 0 or 1"
#endif
#ifndef LSAMD_ABI_DTYPE
#error "LSAMD_ABI_DTYPE must be float or double"
#endif

typedef LSAMD_ABI_DTYPE real_t;

namespace {

constexpr int kVariant = LSAMD_ABI_ERROR ? LSAMD_GENERALIZED : LSAMD_STANDARD;
constexpr int kDtype = sizeof(real_t) == 8 ? LSAMD_F64 : LSAMD_F32;
constexpr const char *kDtypeName = sizeof(real_t) == 8 ? "double" : "float";

int env_form() {
    const char *v = std::getenv("LSAMD_FORM");
    if (v && (std::strcmp(v, "split") == 0 || std::strcmp(v, "1") == 0)) return LSAMD_FORM_SPLIT;
    if (v && (std::strcmp(v, "split32") == 0 || std::strcmp(v, "2") == 0)) return LSAMD_FORM_SPLIT32;
    return LSAMD_FORM_PLAIN;
}

bool env_flag(const char *name, bool dflt) {
    const char *v = std::getenv(name);
    if (!v || !*v) return dflt;
    return !(std::strcmp(v, "0") == 0 || std::strcmp(v, "false") == 0 || std::strcmp(v, "no") == 0);
}

// computePeriod (main.cu:1545-1563) over one object's pgram: period of the
// first strict maximum from power 0.
double cpu_period(const real_t *pg, uint32_t nf, real_t fmin, real_t fstep) {
    real_t best = 0;
    uint32_t idx = 0;
    for (uint32_t i = 0; i < nf; ++i)
        if (best < pg[i]) {
            best = pg[i];
            idx = i;
        }
    return (double)(real_t)((1.0 / (double)(fmin + ((real_t)idx * fstep))) * 2.0 * M_PI);
}

}  // namespace

extern "C" void LombScarglePy(unsigned int *objectId, real_t *timeX, real_t *magY, real_t *magDY,
                              unsigned int sizeData, double minFreq, double maxFreq,
                              unsigned int freqToTest, int MODE, real_t *pgram) {
    std::printf("\nMinimum Frequency: %f", minFreq);
    std::printf("\nMaximum Frequency: %f", maxFreq);
    std::printf("\nNumber of frequencies to test: %u", freqToTest);
    std::printf("\nMode: %d", MODE);
    if (LSAMD_ABI_ERROR)
        std::printf("\nExecuting L-S variant from AstroPy that propogates error and floats the mean");

    const auto t0 = std::chrono::steady_clock::now();
    int rc = LSAMD_OK;
    double sumPeriods = 0.0;
    const int form = env_form();
    const int normalize = LSAMD_ABI_ERROR ? 0 : 1;
    const bool gpu = MODE == 1 || MODE == 2;

    if (gpu) {
        const uint32_t K = MODE == 1 ? lsamd_count_objects(objectId, sizeData) : 1u;
        std::vector<real_t> periods(K > 0 ? K : 1);
        uint32_t kout = 0;
        rc = lsamd_batch_host(kVariant, kDtype, form, normalize, 1, MODE == 1 ? objectId : nullptr,
                              timeX, magY, magDY, sizeData, minFreq, maxFreq, freqToTest, pgram, 1,
                              nullptr, nullptr, periods.data(), &kout);
        if (rc == LSAMD_OK) {
            for (uint32_t k = 0; k < kout; ++k) sumPeriods += (double)periods[k];
            std::printf("\nNumber of unique objects: %u", kout);
        }
    } else if (MODE == 4 || MODE == 5) {
        const int nthreads = (int)std::strtol(std::getenv("LSAMD_CPU_THREADS") ? std::getenv("LSAMD_CPU_THREADS") : "0", nullptr, 10);
        if (MODE == 4)
            rc = lsamd_cpu_batch(kVariant, kDtype, normalize, objectId, timeX, magY, magDY, sizeData,
                                 minFreq, maxFreq, freqToTest, pgram, nthreads);
        else
            rc = lsamd_cpu_single(kVariant, kDtype, normalize, timeX, magY, magDY, sizeData, minFreq,
                                  maxFreq, freqToTest, pgram, nthreads);
        if (rc == LSAMD_OK) {
            const uint32_t K = MODE == 4 ? lsamd_count_objects(objectId, sizeData) : 1u;
            const real_t fmin = (real_t)minFreq, fmax = (real_t)maxFreq;
            const real_t fstep = (real_t)((double)(fmax - fmin) / ((double)freqToTest * 1.0));
            for (uint32_t k = 0; k < K; ++k)
                sumPeriods += cpu_period(pgram + (uint64_t)k * freqToTest, freqToTest, fmin, fstep);
        }
    } else {
        std::fprintf(stderr, "\n[lsamd] LombScarglePy: MODE %d is not 1, 2, 4 or 5; nothing computed\n",
                     MODE);
        std::fflush(stdout);
        return;
    }
    const double secs =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();

    if (rc != LSAMD_OK) {
        std::fflush(stdout);
        std::fprintf(stderr, "\n[lsamd] LombScarglePy failed (%d): %s\n", rc, lsamd_last_error());
        if (env_flag("LSAMD_EXIT_ON_ERROR", false)) std::exit(rc);
        return;
    }
    std::printf("\nTotal time to compute %s: %f", gpu ? "batch" : "pgram", secs);
    if (MODE == 1 || MODE == 4)
        std::printf("\n[Validation] Sum of all periods: %f", sumPeriods);
    else
        std::printf("\n[Validation] Period: %f", sumPeriods);
    std::fflush(stdout);

    if (env_flag("LSAMD_GPU_STATS", true)) {
        if (FILE *f = std::fopen("gpu_stats.txt", "a")) {
            int T = 0, M = 0;
            lsamd_get_tile(form, kVariant, kDtype, &T, &M);
            std::fprintf(f,
                         "%g, Sum of periods: %g, Min/Max Freq: %g/%g,  Num tested freq: %u, MODE: %d, "
                         "ERROR/DTYPE/FORM/TILE: %d, %s, %s, %dx%d\n",
                         secs, sumPeriods, minFreq, maxFreq, freqToTest, MODE, LSAMD_ABI_ERROR,
                         kDtypeName,
                         form == LSAMD_FORM_SPLIT ? "split" : (form == LSAMD_FORM_SPLIT32 ? "split32" : "plain"),
                         T, M);
            std::fclose(f);
        }
    }
}
