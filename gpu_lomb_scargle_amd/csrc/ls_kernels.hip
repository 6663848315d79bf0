// HIP kernels of the MI355X-native Lomb-Scargle sweep and the stream-ordered
// device entry point lsamd_sweep().
//
// Work decomposition: work items (object, frequency chunk), one workgroup
// each, or taken in groups of consecutive chunks.  A chunk is T*M
// consecutive frequency bins; thread tid owns bins chunk*T*M + m*T + tid
// (m < M).  The object's observations are staged through LDS T at a time and
// read back as wave-wide broadcasts; all sums live in registers.
// Per-object normalisation and the argmax/top-3 peak search are fused: every
// workgroup writes its chunk's best bin (or best three) and a small
// reduction kernel merges the chunks of each object.
//
// Three arithmetic forms:
//   plain   - the reference's arithmetic (source/main.cu:1566-1599 standard,
//             main.cu:1701-1777 generalized): DTYPE phase omega*t, accurate
//             sin/cos, sequential DTYPE sums with no FMA contraction, the
//             tau form of the power.  Reproduces the reference CPU path.
//   split   - FP64 phase reduction (revolutions, rounded to nearest), FP32
//             hardware sin/cos seed corrected in FP64, then the thread's M
//             bins by an FP64 rotation and three-term recurrence (step T
//             bins) with FP64 sums.
//   split32 - the same seed in FP32, an FP32 rotation across the bins, and
//             FP32 partial sums over 32-64 observations folded into FP64.
// Both split forms use the tau-free FP64 power
//     standard:    P = (SS XC^2 - 2 CS XC XS + CC XS^2) / (2 (CC SS - CS^2))
//     generalized: P = (SS YC^2 - 2 CS YC YS + CC YS^2) / ((CC SS - CS^2) YY)
// (centred weighted sums), equal to the reference's tau form because the
// 2x2 least-squares projection is rotation invariant.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>

#include "ls_internal.h"

namespace lsamd {

namespace {

// Observations per FP32 partial sum of the split32 sweep before it is folded
// into FP64.  Standard 32: on the 1000-object batch the worst relative power
// error drops from 9.4e-6 (64) to 7.0e-6 for ~1% of the time.  Generalized
// 64: even 16 stays above 1e-5 there (1.2e-5), at ~7% of the time.
constexpr uint32_t kFlushStd = 32, kFlushGen = 64;
constexpr uint32_t kGroup = 4;  // work items per workgroup fetch, split forms (a power of two)
constexpr uint64_t kGroupMinItems = 65536;  // fewer items: one per fetch (e.g. a single object)
// Split forms: below this many (object, chunk) items, long objects are also
// split by observations (at least T observations per slice), up to about
// this many items.
constexpr uint64_t kSliceTargetItems = 2048;
constexpr int kPeaks = 3;      // peaks tracked per object (top-1 and top-3)
constexpr int kGen = 8;        // doubles per object of split-generalized constants
constexpr double kInv2Pi = 0.159154943091895335768883763372514362;  // 1/(2 pi)
constexpr double kTwoPi = 6.283185307179586476925286766559005768;   // 2 pi

// ---------------------------------------------------------------------------
// Peak lists: best kPeaks (power, bin) pairs, ordered by power descending and
// bin ascending.  Initialised to (0, 0): a bin enters only with power > 0,
// and an equal power never displaces an earlier bin -- the reference's
// "first index of the strict maximum, starting from 0" (main.cu:1140-1149,
// 1545-1556) and its three-scan top-3 (main.cu:1165-1219).  NaN never enters.
// ---------------------------------------------------------------------------
struct Peaks {
    double p[kPeaks];
    uint32_t i[kPeaks];
};

__device__ __forceinline__ bool better(double p1, uint32_t i1, double p2, uint32_t i2) {
    return p1 > p2 || (p1 == p2 && i1 < i2);
}

__device__ __forceinline__ void peaks_init(Peaks &L) {
#pragma unroll
    for (int r = 0; r < kPeaks; ++r) {
        L.p[r] = 0.0;
        L.i[r] = 0u;
    }
}

// Insertion by compare-and-swap down the list (static indices only).
__device__ __forceinline__ void peaks_insert(Peaks &L, double p, uint32_t i) {
#pragma unroll
    for (int r = 0; r < kPeaks; ++r) {
        if (better(p, i, L.p[r], L.i[r])) {
            const double tp = L.p[r];
            const uint32_t ti = L.i[r];
            L.p[r] = p;
            L.i[r] = i;
            p = tp;
            i = ti;
        }
    }
}

// Adds bin i to the list kept for top_k = 1 (the best bin only) or 3;
// nothing for top_k = 0.  Each thread adds its bins in ascending order.
__device__ __forceinline__ void peaks_add(Peaks &L, int top_k, double p, uint32_t i) {
    if (top_k == 1) {
        if (better(p, i, L.p[0], L.i[0])) {
            L.p[0] = p;
            L.i[0] = i;
        }
    } else if (top_k) {
        peaks_insert(L, p, i);
    }
}

// Best (power, bin) of a whole workgroup; the result is valid in thread 0.
template <int T>
__device__ __forceinline__ void block_best(double &p, uint32_t &i, double (*s_p)[kPeaks],
                                           uint32_t (*s_i)[kPeaks]) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const double op = __shfl_xor(p, off, 64);
        const uint32_t oi = __shfl_xor(i, off, 64);
        if (better(op, oi, p, i)) {
            p = op;
            i = oi;
        }
    }
    if constexpr (T > 64) {
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        if (lane == 0) {
            s_p[wave][0] = p;
            s_i[wave][0] = i;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int w = 1; w < T / 64; ++w)
                if (better(s_p[w][0], s_i[w][0], p, i)) {
                    p = s_p[w][0];
                    i = s_i[w][0];
                }
        }
    }
}

// Merge the lists of a whole workgroup; the result is valid in thread 0.
template <int T>
__device__ __forceinline__ void block_peaks(Peaks &L, double (*s_p)[kPeaks],
                                            uint32_t (*s_i)[kPeaks]) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        double op[kPeaks];
        uint32_t oi[kPeaks];
#pragma unroll
        for (int r = 0; r < kPeaks; ++r) {
            op[r] = __shfl_xor(L.p[r], off, 64);
            oi[r] = __shfl_xor(L.i[r], off, 64);
        }
#pragma unroll
        for (int r = 0; r < kPeaks; ++r) peaks_insert(L, op[r], oi[r]);
    }
    if constexpr (T > 64) {
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        if (lane == 0) {
#pragma unroll
            for (int r = 0; r < kPeaks; ++r) {
                s_p[wave][r] = L.p[r];
                s_i[wave][r] = L.i[r];
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int w = 1; w < T / 64; ++w)
#pragma unroll
                for (int r = 0; r < kPeaks; ++r) peaks_insert(L, s_p[w][r], s_i[w][r]);
        }
    }
}

// ---------------------------------------------------------------------------
// Reference-arithmetic helpers (plain form)
// ---------------------------------------------------------------------------
// The FP32 plain form evaluates sin, cos and atan2 of its float arguments in
// FP64 and rounds once: correctly rounded float results (double rounding is
// possible only within 2^-29 of a float midpoint).  The reference's float CPU
// build calls glibc's sinf/cosf/atan2f, which are correctly rounded on most
// but not all inputs, so this is the closest reproducible target.
__device__ __forceinline__ void sincos_r(double x, double *s, double *c) { sincos(x, s, c); }
__device__ __forceinline__ void sincos_r(float x, float *s, float *c) {
    double sd, cd;
    sincos((double)x, &sd, &cd);
    *s = (float)sd;
    *c = (float)cd;
}
__device__ __forceinline__ double atan2_r(double y, double x) { return atan2(y, x); }
__device__ __forceinline__ float atan2_r(float y, float x) { return (float)atan2((double)y, (double)x); }
__device__ __forceinline__ double sqrt_r(double x) { return sqrt(x); }
__device__ __forceinline__ float sqrt_r(float x) { return sqrtf(x); }

// ---------------------------------------------------------------------------
// Kernel arguments
// ---------------------------------------------------------------------------
struct StatsArgs {
    const uint64_t *offsets;
    const void *y, *dy;
    uint32_t K;
    int normalize;       // scale = 2/(N sigma^2)  (main.cu:1295-1352)
    int precenter;       // plain generalized: shift = sum(w y)/sum(w), w=1/sqrt(dy)
    int split_gen;       // split generalized constants
    double *obj_shift, *obj_scale, *obj_gen;
};

struct RecArgs {
    const uint64_t *offsets;
    const void *t, *y, *dy;
    const double *obj_shift, *obj_gen;
    void *rec_y, *rec_w;       // plain
    double *rec_t;             // split
    float4 *rec_a;             // split: {y or sqrt(w)(y-ybar), -sin step or sqrt(w), cos step, sin step} (FP32)
    float *rec_n;              // split generalized: -sin step (FP32)
    double2 *rec_r;            // split: {cos step, sin step} (FP64)
    double step_rev;           // split: T * df (revolutions/day per T bins)
    int precenter;
};

struct SweepArgs {
    const uint64_t *offsets;
    const uint32_t *order;
    uint32_t K, nf, chunks;
    uint64_t item_base;        // first item of this launch
    uint64_t total_items;      // K * chunks
    uint32_t group;            // consecutive items per workgroup fetch (a power of two)
    uint32_t slices;           // split forms: observation slices per (object, chunk), >= 1
    double *part_sums;         // slices > 1: per-slice FP64 sums [K*chunks][slices][NS][M][T]
    unsigned long long *queue; // persistent schedule: next item counter, else null
    // plain
    const void *t, *y, *w;
    double fmin_r, fstep_r;    // DTYPE values stored in double
    // split
    const double *rec_t;
    const float4 *rec_a;
    const float *rec_n;
    const double2 *rec_r;
    double fmin_rev, df_rev;   // revolutions/day
    const double *obj_gen;
    // common
    const double *obj_scale;
    void *pgram;
    double *part_pow;
    uint32_t *part_idx;
    int top_k;                 // peaks kept per chunk: 0, 1 or 3
};

struct ReduceArgs {
    uint32_t K, chunks, top_k;
    const double *part_pow;
    const uint32_t *part_idx;
    uint32_t *peak_index;
    void *peak_power, *peak_period;
    double pmin, pstep;        // period of bin i = 2 pi / (pmin + i*pstep)
};

// ---------------------------------------------------------------------------
// Per-object statistics: one thread per object, sums in index order so that
// the values equal the CPU path's sequential sums.
// ---------------------------------------------------------------------------
template <typename R>
__global__ void k_obj_stats(StatsArgs a) {
#pragma clang fp contract(off)
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.K) return;
    const uint64_t b = a.offsets[k], e = a.offsets[k + 1];
    const R *y = static_cast<const R *>(a.y);
    const R *dy = static_cast<const R *>(a.dy);
    const double nd = (double)(e - b);
    double scale = 1.0;
    if (a.normalize) {  // computeStandardDevSquared, main.cu:1295-1317
        double sum = 0.0;
        for (uint64_t i = b; i < e; ++i) sum += (double)y[i];
        const double mean = sum / (nd * 1.0);
        double sum2 = 0.0;
        for (uint64_t i = b; i < e; ++i) sum2 += ((double)y[i] - mean) * ((double)y[i] - mean);
        const double sigma = sqrt(sum2 / (nd * 1.0));
        scale = 2.0 / (nd * (sigma * sigma));
    }
    a.obj_scale[k] = scale;
    if (a.precenter) {  // updateYerrorfactor, main.cu:1631-1662 (index order)
        R sumw = 0, dot = 0;
        for (uint64_t i = b; i < e; ++i) sumw += (R)(1.0 / (double)sqrt_r(dy[i]));
        for (uint64_t i = b; i < e; ++i) dot += (R)(1.0 / (double)sqrt_r(dy[i])) * y[i];
        a.obj_shift[k] = (double)(dot / sumw);
    }
    if (a.split_gen) {
        // weights 1/dy^2 normalised to unit sum; y centred on its weighted
        // mean (the generalized power does not depend on the shift).  The
        // sweep carries sqrt(w) in each observation's phasor, so the weights
        // it applies are exactly qf^2 with qf = (float)sqrt(w): the same
        // values are summed here (qf^2 is exact in FP64).
        double W = 0.0, WY = 0.0;
        for (uint64_t i = b; i < e; ++i) {
            const double wi = 1.0 / ((double)dy[i] * (double)dy[i]);
            W += wi;
            WY += wi * (double)y[i];
        }
        const double ybar = WY / W;
        double Wf = 0.0, Ys = 0.0, YYs = 0.0;
        for (uint64_t i = b; i < e; ++i) {
            const double qi = sqrt(1.0 / ((double)dy[i] * (double)dy[i]) / W);
            const double yc = (double)y[i] - ybar;
            const double q = (double)(float)qi;
            const float bq = (float)(qi * yc);
            Wf += q * q;
            Ys += (double)bq * q;
            YYs += q * q * yc * yc;
        }
        const double Yb = Ys / Wf;
        a.obj_gen[kGen * k + 0] = Wf;
        a.obj_gen[kGen * k + 1] = Yb;
        a.obj_gen[kGen * k + 2] = YYs / Wf - Yb * Yb;
        a.obj_gen[kGen * k + 3] = ybar;
        a.obj_gen[kGen * k + 4] = W;
    }
}

__device__ __forceinline__ double wave_sum64(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// The same statistics for the split forms, one wavefront per object: sums in
// lane-strided order with a butterfly reduction (the split forms are not
// bit-matched to the CPU path, and an object's sums no longer wait on one
// thread's sequential loop over its observations).
template <typename R>
__global__ __launch_bounds__(64) void k_obj_stats_wave(StatsArgs a) {
    const uint32_t k = blockIdx.x;
    const uint32_t lane = threadIdx.x;
    const uint64_t b = a.offsets[k], e = a.offsets[k + 1];
    const R *y = static_cast<const R *>(a.y);
    const R *dy = static_cast<const R *>(a.dy);
    const double nd = (double)(e - b);
    double scale = 1.0;
    if (a.normalize) {  // 2/(N sigma^2), sigma^2 the population variance
        double s = 0.0;
        for (uint64_t i = b + lane; i < e; i += 64) s += (double)y[i];
        const double mean = wave_sum64(s) / nd;
        double s2 = 0.0;
        for (uint64_t i = b + lane; i < e; i += 64) s2 += ((double)y[i] - mean) * ((double)y[i] - mean);
        const double sigma2 = wave_sum64(s2) / nd;
        scale = 2.0 / (nd * sigma2);
    }
    if (lane == 0) a.obj_scale[k] = scale;
    if (a.split_gen) {  // as k_obj_stats
        double W = 0.0, WY = 0.0;
        for (uint64_t i = b + lane; i < e; i += 64) {
            const double wi = 1.0 / ((double)dy[i] * (double)dy[i]);
            W += wi;
            WY += wi * (double)y[i];
        }
        W = wave_sum64(W);
        const double ybar = wave_sum64(WY) / W;
        double Wf = 0.0, Ys = 0.0, YYs = 0.0;
        for (uint64_t i = b + lane; i < e; i += 64) {
            const double qi = sqrt(1.0 / ((double)dy[i] * (double)dy[i]) / W);
            const double yc = (double)y[i] - ybar;
            const double q = (double)(float)qi;
            const float bq = (float)(qi * yc);
            Wf += q * q;
            Ys += (double)bq * q;
            YYs += q * q * yc * yc;
        }
        Wf = wave_sum64(Wf);
        Ys = wave_sum64(Ys);
        YYs = wave_sum64(YYs);
        if (lane == 0) {
            const double Yb = Ys / Wf;
            a.obj_gen[kGen * k + 0] = Wf;
            a.obj_gen[kGen * k + 1] = Yb;
            a.obj_gen[kGen * k + 2] = YYs / Wf - Yb * Yb;
            a.obj_gen[kGen * k + 3] = ybar;
            a.obj_gen[kGen * k + 4] = W;
        }
    }
}

// ---------------------------------------------------------------------------
// Per-observation records.  One workgroup per object.
// ---------------------------------------------------------------------------
template <typename R, bool GEN>
__global__ void k_records_plain(RecArgs a) {
#pragma clang fp contract(off)
    const uint32_t k = blockIdx.x;
    const uint64_t b = a.offsets[k], e = a.offsets[k + 1];
    const R *y = static_cast<const R *>(a.y);
    const R *dy = static_cast<const R *>(a.dy);
    R *ry = static_cast<R *>(a.rec_y);
    R *rw = static_cast<R *>(a.rec_w);
    const R shift = a.precenter ? (R)a.obj_shift[k] : (R)0;
    for (uint64_t i = b + threadIdx.x; i < e; i += blockDim.x) {
        if (a.precenter) ry[i] = y[i] - shift;  // main.cu:1656
        if (GEN) {
            R w = (R)(1.0 / (double)dy[i]);  // main.cu:1713-1714
            w *= w;
            rw[i] = w;
        }
    }
}

template <typename R, bool GEN>
__global__ void k_records_split(RecArgs a) {
    const uint32_t k = blockIdx.x;
    const uint64_t b = a.offsets[k], e = a.offsets[k + 1];
    const R *t = static_cast<const R *>(a.t);
    const R *y = static_cast<const R *>(a.y);
    const R *dy = static_cast<const R *>(a.dy);
    double W = 1.0, ybar = 0.0;
    if (GEN) {  // the same W and ybar k_obj_stats used, so the float weights match
        ybar = a.obj_gen[kGen * k + 3];
        W = a.obj_gen[kGen * k + 4];
    }
    for (uint64_t i = b + threadIdx.x; i < e; i += blockDim.x) {
        const double ti = (double)t[i];
        double ps = a.step_rev * ti;
        ps -= rint(ps);
        double sv, cv;
        sincospi(2.0 * ps, &sv, &cv);
        // The FP32 rotation needs both sin D and -sin D as plain operands (an
        // all-VGPR v_fmac_f32 issues at twice the rate of a negated VOP3 form).
        float4 r;
        if (GEN) {
            // {sqrt(w) (y - ybar), sqrt(w)}: the sweep scales the phasor by sqrt(w)
            const double qi = sqrt(1.0 / ((double)dy[i] * (double)dy[i]) / W);
            r.x = (float)(qi * ((double)y[i] - ybar));
            r.y = (float)qi;
            a.rec_n[i] = -(float)sv;
        } else {
            r.x = (float)y[i];
            r.y = -(float)sv;
        }
        r.z = (float)cv;
        r.w = (float)sv;
        a.rec_t[i] = ti;
        a.rec_a[i] = r;
        a.rec_r[i] = make_double2(cv, sv);
    }
}

__device__ __forceinline__ void item_coords(const SweepArgs &a, uint64_t item, uint32_t &obj,
                                            uint32_t &chunk) {
    const uint32_t pos = (uint32_t)(item / a.chunks);
    chunk = (uint32_t)(item - (uint64_t)pos * a.chunks);
    obj = a.order ? a.order[pos] : pos;
}

// Split forms: item = (object position, chunk, observation slice), slice
// minor.  With one slice an item is the whole object; with S slices it is
// observations [beg + s L, beg + (s + 1) L) of the object, L = ceil(N / S).
__device__ __forceinline__ void item_coords_sliced(const SweepArgs &a, uint64_t item, uint32_t &obj,
                                                   uint32_t &chunk, uint32_t &slice, uint64_t &beg,
                                                   uint64_t &end) {
    uint64_t oc = item;
    slice = 0;
    if (a.slices > 1) {
        oc = item / a.slices;
        slice = (uint32_t)(item - oc * a.slices);
    }
    item_coords(a, oc, obj, chunk);
    beg = a.offsets[obj];
    end = a.offsets[obj + 1];
    if (a.slices > 1) {
        const uint64_t len = (end - beg + a.slices - 1) / a.slices;
        const uint64_t b = beg + (uint64_t)slice * len;
        beg = b < end ? b : end;
        end = (beg + len) < end ? beg + len : end;
    }
}

// Work items are numbered in processing order (longest object first),
// chunk-minor, and taken in aligned groups of a.group consecutive items (the
// next chunks of the same object, mostly): either one group per workgroup
// (item_base + blockIdx.x * group), or, with a queue, persistent workgroups
// that take the next group from a counter until the items run out.
__device__ __forceinline__ uint64_t fetch_item(const SweepArgs &a, uint64_t *s_item) {
    __syncthreads();  // every thread is done with the previous item and its s_item read
    if (threadIdx.x == 0) *s_item = atomicAdd(a.queue, (unsigned long long)a.group);
    __syncthreads();
    return *s_item;
}
__device__ __forceinline__ uint64_t first_item(const SweepArgs &a, uint64_t *s_item) {
    return a.queue ? fetch_item(a, s_item) : a.item_base + (uint64_t)blockIdx.x * a.group;
}
__device__ __forceinline__ uint64_t next_item(const SweepArgs &a, uint64_t *s_item, uint64_t item) {
    const uint64_t n = item + 1;
    if (((n - a.item_base) & (uint64_t)(a.group - 1)) != 0) return n;  // same group
    return a.queue ? fetch_item(a, s_item) : ~0ull;
}

__global__ void k_zero_counter(unsigned long long *c) { *c = 0ull; }

template <typename OUT, int T>
__device__ __forceinline__ void write_partial(const SweepArgs &a, Peaks &L, uint32_t obj,
                                              uint32_t chunk, double (*s_p)[kPeaks],
                                              uint32_t (*s_i)[kPeaks]) {
    const uint64_t o = ((uint64_t)obj * a.chunks + chunk) * kPeaks;
    if (a.top_k == 1) {  // entries 1 and 2 of the chunk's slot stay unused
        block_best<T>(L.p[0], L.i[0], s_p, s_i);
        if (threadIdx.x == 0) {
            a.part_pow[o] = L.p[0];
            a.part_idx[o] = L.i[0];
        }
        return;
    }
    block_peaks<T>(L, s_p, s_i);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int r = 0; r < kPeaks; ++r) {
            a.part_pow[o + r] = L.p[r];
            a.part_idx[o + r] = L.i[r];
        }
    }
}

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
// bin, one hardware sin/cos pair, then the other M-1 bins by FP32 rotation
// up and down with the observation's precomputed step (cos, sin of T bins);
// FP32 partial sums over 32 (standard) or 64 (generalized) observations are
// folded into FP64 accumulators.  Every FP32 instruction of the inner loop is an all-VGPR v_fmac_f32 /
// v_mul_f32 (the records carry -sin D so that no negated operand forces the
// slower VOP3 encoding).
// ---------------------------------------------------------------------------
template <bool GEN, typename OUT, int T, int M>
__global__ __launch_bounds__(T) void k_sweep_split32(SweepArgs a) {
    // standard: XC, XS, CC, CS (SS = N - CC); generalized: all seven sums
    // (its centred SS - S^2 cancels, so SS is accumulated, not derived)
    constexpr int NS = GEN ? 7 : 4;
    constexpr int kSeed = (M - 1) / 2;  // the bin whose phasor is evaluated directly
    constexpr uint32_t kF = GEN ? kFlushGen : kFlushStd;
    __shared__ double s_t[T];
    __shared__ float4 s_a[T];
    __shared__ float s_n[GEN ? T : 1];
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
            if (!reuse) {
                if (tid < cnt) {
                    s_t[tid] = a.rec_t[b0 + tid];
                    s_a[tid] = a.rec_a[b0 + tid];
                    if constexpr (GEN) s_n[tid] = a.rec_n[b0 + tid];
                }
                __syncthreads();
            }
            for (uint32_t j0 = 0; j0 < cnt; j0 += kF) {
                const uint32_t j1 = (j0 + kF < cnt) ? j0 + kF : cnt;
                float p[NS][M];
    #pragma unroll
                for (int q = 0; q < NS; ++q)
    #pragma unroll
                    for (int m = 0; m < M; ++m) p[q][m] = 0.0f;
    #pragma unroll 2
                for (uint32_t j = j0; j < j1; ++j) {
                    const double tj = s_t[j];
                    const float4 r = s_a[j];
                    const float nsd = GEN ? s_n[j] : r.y;  // -sin D
                    double ph = fseed * tj;
                    ph -= __builtin_rint(ph);
                    const float phf = (float)ph;
                    float c = __builtin_amdgcn_cosf(phf);
                    float s = __builtin_amdgcn_sinf(phf);
                    if constexpr (GEN) {
                        // u = sqrt(w) z: the rotation is linear, and every
                        // weighted sum becomes one FMA (w c = sqrt(w) u_c,
                        // w c^2 = u_c^2, w y c = (sqrt(w) y) u_c)
                        c *= r.y;
                        s *= r.y;
                    }
                    auto add = [&](int m, float cv, float sv) {
                        if constexpr (GEN) {
                            p[0][m] = fmaf(r.y, cv, p[0][m]);
                            p[1][m] = fmaf(r.y, sv, p[1][m]);
                            p[2][m] = fmaf(r.x, cv, p[2][m]);
                            p[3][m] = fmaf(r.x, sv, p[3][m]);
                            p[4][m] = fmaf(cv, cv, p[4][m]);
                            p[5][m] = fmaf(sv, sv, p[5][m]);
                            p[6][m] = fmaf(cv, sv, p[6][m]);
                        } else {
                            p[0][m] = fmaf(r.x, cv, p[0][m]);
                            p[1][m] = fmaf(r.x, sv, p[1][m]);
                            p[2][m] = fmaf(cv, cv, p[2][m]);
                            p[3][m] = fmaf(cv, sv, p[3][m]);
                        }
                    };
                    // The seed sits at bin kSeed; the rotation walks up to
                    // bin M-1 and down to bin 0, so no bin is more than
                    // M/2 rotations from its seed.
                    add(kSeed, c, s);
                    {
                        float cu = c, su = s;
    #pragma unroll
                        for (int m = kSeed + 1; m < M; ++m) {
                            const float cn = fmaf(cu, r.z, su * nsd);
                            const float sn = fmaf(su, r.z, cu * r.w);
                            cu = cn;
                            su = sn;
                            add(m, cu, su);
                        }
                    }
    #pragma unroll
                    for (int m = kSeed - 1; m >= 0; --m) {
                        const float cn = fmaf(c, r.z, s * r.w);
                        const float sn = fmaf(s, r.z, c * nsd);
                        c = cn;
                        s = sn;
                        add(m, c, s);
                    }
                }
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
// Merge the chunk partials of each object; period of bin i per the GPU host
// path's formula (main.cu:1152).
// ---------------------------------------------------------------------------
// One wave per object: lanes merge the chunk partials at a stride of 64,
// then a shuffle butterfly merges the lanes; lane 0 writes.
template <typename OUT>
__global__ __launch_bounds__(64) void k_reduce_peaks(ReduceArgs a) {
#pragma clang fp contract(off)
    const uint32_t k = blockIdx.x;
    Peaks L;
    peaks_init(L);
    const uint64_t base = (uint64_t)k * a.chunks * kPeaks;
    const int nr = a.top_k == 1 ? 1 : kPeaks;  // top-1 sweeps write entry 0 of each slot
    for (uint32_t c = threadIdx.x; c < a.chunks; c += 64)
        for (int r = 0; r < nr; ++r)
            peaks_insert(L, a.part_pow[base + (uint64_t)c * kPeaks + r],
                         a.part_idx[base + (uint64_t)c * kPeaks + r]);
    block_peaks<64>(L, nullptr, nullptr);
    if (threadIdx.x != 0) return;
    OUT *pp = static_cast<OUT *>(a.peak_power);
    OUT *pe = static_cast<OUT *>(a.peak_period);
    for (uint32_t r = 0; r < a.top_k && r < (uint32_t)kPeaks; ++r) {
        const uint64_t o = (uint64_t)k * a.top_k + r;
        if (a.peak_index) a.peak_index[o] = L.i[r];
        if (pp) pp[o] = (OUT)L.p[r];
        if (pe) pe[o] = (OUT)((1.0 / (a.pmin + (double)L.i[r] * a.pstep)) * 2.0 * M_PI);
    }
}

// ---------------------------------------------------------------------------
// Tile registry
// ---------------------------------------------------------------------------
typedef void (*SweepFn)(SweepArgs);

struct TileEntry {
    int form, variant, dtype, T, M;
    SweepFn fn;
    SweepFn combine;  // split forms: sums the observation slices (null: plain)
};

#define LS_SPLIT32(GEN, OUTT, DT, TT, MM)                                      \
    {LSAMD_FORM_SPLIT32, GEN, DT, TT, MM, k_sweep_split32<(GEN) != 0, OUTT, TT, MM>, \
     k_combine_slices<true, (GEN) != 0, OUTT, TT, MM>}
#define LS_SPLIT(GEN, OUTT, DT, TT, MM)                                      \
    {LSAMD_FORM_SPLIT, GEN, DT, TT, MM, k_sweep_split64<(GEN) != 0, OUTT, TT, MM>, \
     k_combine_slices<false, (GEN) != 0, OUTT, TT, MM>}
#define LS_PLAIN_STD(R, DT, TT, MM) \
    {LSAMD_FORM_PLAIN, 0, DT, TT, MM, k_sweep_plain_std<R, TT, MM>, nullptr}
#define LS_PLAIN_GEN(R, DT, TT, MM) \
    {LSAMD_FORM_PLAIN, 1, DT, TT, MM, k_sweep_plain_gen<R, TT, MM>, nullptr}

const TileEntry kTiles[] = {
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
    // plain form
    LS_PLAIN_STD(float, LSAMD_F32, 256, 1), LS_PLAIN_STD(double, LSAMD_F64, 256, 1),
    LS_PLAIN_STD(float, LSAMD_F32, 256, 2), LS_PLAIN_STD(double, LSAMD_F64, 256, 2),
    LS_PLAIN_GEN(float, LSAMD_F32, 256, 1), LS_PLAIN_GEN(double, LSAMD_F64, 256, 1),
};

const TileEntry *find_tile(int form, int variant, int dtype, int T, int M) {
    for (const TileEntry &e : kTiles)
        if (e.form == form && e.variant == variant && e.dtype == dtype && e.T == T && e.M == M)
            return &e;
    return nullptr;
}

// user overrides: [form][variant] -> (T, M); 0 = built-in
constexpr int kForms = 3;
int g_tile_T[kForms][2] = {{0, 0}, {0, 0}, {0, 0}};
int g_tile_M[kForms][2] = {{0, 0}, {0, 0}, {0, 0}};

// Built-in shapes: the fastest of the tile sweep on the 1000-object x 1e6
// batches (MI355X): split 256x16 (standard) / 256x8 (generalized); split32
// 256x8 (standard) / 256x4 (generalized).  lsamd_autotune re-picks per
// workload.
Tile default_tile(int form, int variant) {
    if (form == LSAMD_FORM_SPLIT) return variant ? Tile{256, 8} : Tile{256, 16};
    if (form == LSAMD_FORM_SPLIT32) return variant ? Tile{256, 4} : Tile{256, 8};
    return Tile{256, 1};
}

bool g_timing = false;
double g_last_ms = -1.0;
hipEvent_t g_ev[2] = {nullptr, nullptr};
bool g_ev_pending = false;

size_t align_up(size_t x) { return (x + 255) & ~(size_t)255; }

}  // namespace

Tile current_tile(int form, int variant, int dtype) {
    (void)dtype;
    if (form < 0 || form >= kForms || variant < 0 || variant > 1) return Tile{0, 0};
    if (g_tile_T[form][variant] > 0) return Tile{g_tile_T[form][variant], g_tile_M[form][variant]};
    return default_tile(form, variant);
}

bool tile_supported(int form, int variant, int threads, int fpt) {
    return find_tile(form, variant, LSAMD_F32, threads, fpt) != nullptr &&
           find_tile(form, variant, LSAMD_F64, threads, fpt) != nullptr;
}

int plan_workspace(const lsamd_sweep_args *a, Workspace *w) {
    *w = Workspace{};
    const Tile tile = current_tile(a->form, a->variant, a->dtype);
    w->tile = tile;
    const uint64_t per_chunk = (uint64_t)tile.threads * tile.fpt;
    w->chunks = (uint32_t)((a->num_freqs + per_chunk - 1) / per_chunk);
    const size_t K = a->num_objects, n = a->num_obs;
    const size_t rsz = a->dtype == LSAMD_F64 ? 8 : 4;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t o = off;
        off = align_up(off + bytes);
        return o;
    };
    w->obj_shift = take(8 * K);
    w->obj_scale = take(8 * K);
    w->obj_gen = take(8 * kGen * K);
    if (a->form == LSAMD_FORM_PLAIN) {
        w->rec_y = take(rsz * n);
        w->rec_w = take(rsz * n);
    } else {
        w->rec_t = take(8 * n);
        w->rec_a = take(16 * n);
        w->rec_n = take(4 * n);
        w->rec_r = take(16 * n);
    }
    const size_t parts = (size_t)K * w->chunks * kPeaks;
    w->part_pow = take(8 * parts);
    w->part_idx = take(4 * parts);
    // Observation slices (split forms): a single object whose (object, chunk)
    // items are too few to fill the GPU has its observations split over
    // workgroups whose FP64 sums k_combine_slices adds up.  Only single-object
    // sweeps are sliced, so an object's result never depends on which other
    // objects share its batch.
    w->slices = 1;
    if (a->form != LSAMD_FORM_PLAIN && a->obs_slices == 0 && K == 1) {
        const uint64_t base = (uint64_t)K * w->chunks;
        const uint64_t nbar = (n + K - 1) / K;
        const uint64_t by_items = (kSliceTargetItems + base - 1) / base;
        const uint64_t by_obs = (nbar + tile.threads - 1) / tile.threads;
        const uint64_t sl = std::min(by_items, by_obs);
        if (base < kSliceTargetItems && sl >= 2) {
            w->slices = (uint32_t)sl;
            const int ns = a->variant == LSAMD_GENERALIZED ? 7 : 4;
            w->part_sums = take(8 * base * sl * (size_t)ns * tile.fpt * tile.threads);
        }
    }
    w->queue = take(8);
    w->bytes = off + 256;
    return LSAMD_OK;
}

}  // namespace lsamd

using namespace lsamd;

extern "C" {

size_t lsamd_sweep_workspace_size(const lsamd_sweep_args *a) {
    if (!a) return 0;
    Workspace w;
    plan_workspace(a, &w);
    return w.bytes;
}

int lsamd_set_tile(int form, int variant, int threads, int fpt) {
    if (form < 0 || form >= kForms || variant < 0 || variant > 1)
        return set_error(LSAMD_ERR_INVALID, "lsamd_set_tile: bad form/variant");
    if (threads == 0 || fpt == 0) {
        g_tile_T[form][variant] = g_tile_M[form][variant] = 0;
        return LSAMD_OK;
    }
    if (!tile_supported(form, variant, threads, fpt))
        return set_error(LSAMD_ERR_INVALID, "lsamd_set_tile: tile shape " + std::to_string(threads) +
                                                "x" + std::to_string(fpt) + " not compiled");
    g_tile_T[form][variant] = threads;
    g_tile_M[form][variant] = fpt;
    return LSAMD_OK;
}

void lsamd_get_tile(int form, int variant, int dtype, int *threads, int *fpt) {
    const Tile t = current_tile(form, variant, dtype);
    if (threads) *threads = t.threads;
    if (fpt) *fpt = t.fpt;
}

void lsamd_set_timing(int enabled) { g_timing = enabled != 0; }

double lsamd_last_sweep_kernel_ms(void) {
    if (g_ev_pending) {
        float ms = -1.0f;
        if (hipEventElapsedTime(&ms, g_ev[0], g_ev[1]) == hipSuccess) g_last_ms = ms;
        g_ev_pending = false;
    }
    return g_last_ms;
}

#define LS_HIP(call)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess)                                                                \
            return set_error(LSAMD_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

int lsamd_sweep(const lsamd_sweep_args *a, void *workspace, size_t workspace_bytes, void *stream_v) {
    if (!a) return set_error(LSAMD_ERR_INVALID, "lsamd_sweep: null args");
    if (a->variant != LSAMD_STANDARD && a->variant != LSAMD_GENERALIZED)
        return set_error(LSAMD_ERR_INVALID, "lsamd_sweep: bad variant");
    if (a->dtype != LSAMD_F32 && a->dtype != LSAMD_F64)
        return set_error(LSAMD_ERR_INVALID, "lsamd_sweep: bad dtype");
    if (a->form != LSAMD_FORM_PLAIN && a->form != LSAMD_FORM_SPLIT && a->form != LSAMD_FORM_SPLIT32)
        return set_error(LSAMD_ERR_INVALID, "lsamd_sweep: bad form");
    if (a->num_objects == 0 || a->num_freqs == 0 || a->num_obs == 0)
        return set_error(LSAMD_ERR_INVALID, "lsamd_sweep: empty input (objects, frequencies and "
                                            "observations must all be >= 1)");
    if (!a->offsets || !a->t || !a->y)
        return set_error(LSAMD_ERR_INVALID, "lsamd_sweep: null device pointer");
    if (a->variant == LSAMD_GENERALIZED && !a->dy)
        return set_error(LSAMD_ERR_INVALID, "lsamd_sweep: generalized variant needs dy");
    if (a->top_k != 0 && a->top_k != 1 && a->top_k != 3)
        return set_error(LSAMD_ERR_INVALID, "lsamd_sweep: top_k must be 0, 1 or 3");
    if (!(a->max_freq > a->min_freq))
        return set_error(LSAMD_ERR_INVALID, "lsamd_sweep: max_freq must exceed min_freq");
    if (a->pgram == nullptr && a->top_k == 0)
        return set_error(LSAMD_ERR_INVALID, "lsamd_sweep: nothing to compute (no pgram, top_k=0)");
    if (a->schedule != LSAMD_SCHEDULE_GRID && a->schedule != LSAMD_SCHEDULE_PERSISTENT)
        return set_error(LSAMD_ERR_INVALID, "lsamd_sweep: bad schedule");
    if (a->obs_slices != 0 && a->obs_slices != 1)
        return set_error(LSAMD_ERR_INVALID, "lsamd_sweep: obs_slices must be 0 or 1");
    Workspace w;
    plan_workspace(a, &w);
    if (!workspace || workspace_bytes < w.bytes)
        return set_error(LSAMD_ERR_INVALID, "lsamd_sweep: workspace too small (need " +
                                                std::to_string(w.bytes) + " bytes)");
    const TileEntry *te = find_tile(a->form, a->variant, a->dtype, w.tile.threads, w.tile.fpt);
    if (!te) return set_error(LSAMD_ERR_INVALID, "lsamd_sweep: no kernel for this tile");

    hipStream_t st = static_cast<hipStream_t>(stream_v);
    char *ws = static_cast<char *>(workspace);
    ws = reinterpret_cast<char *>((reinterpret_cast<uintptr_t>(ws) + 255) & ~(uintptr_t)255);
    double *obj_shift = reinterpret_cast<double *>(ws + w.obj_shift);
    double *obj_scale = reinterpret_cast<double *>(ws + w.obj_scale);
    double *obj_gen = reinterpret_cast<double *>(ws + w.obj_gen);
    const bool f64 = a->dtype == LSAMD_F64;
    const bool gen = a->variant == LSAMD_GENERALIZED;
    const bool split = a->form != LSAMD_FORM_PLAIN;
    const bool plain_pre = !split && gen && a->precenter;
    const uint32_t K = a->num_objects;
    const int T = w.tile.threads, M = w.tile.fpt;

    // 1. per-object statistics
    StatsArgs sa{a->offsets, a->y, a->dy, K, a->normalize && !gen, plain_pre ? 1 : 0,
                 (split && gen) ? 1 : 0, obj_shift, obj_scale, obj_gen};
    if (split) {
        if (f64)
            hipLaunchKernelGGL(k_obj_stats_wave<double>, dim3(K), dim3(64), 0, st, sa);
        else
            hipLaunchKernelGGL(k_obj_stats_wave<float>, dim3(K), dim3(64), 0, st, sa);
    } else {
        const dim3 sb(128), sg((K + 127) / 128);
        if (f64)
            hipLaunchKernelGGL(k_obj_stats<double>, sg, sb, 0, st, sa);
        else
            hipLaunchKernelGGL(k_obj_stats<float>, sg, sb, 0, st, sa);
    }
    LS_HIP(hipGetLastError());

    // 2. per-observation records
    RecArgs ra{};
    ra.offsets = a->offsets;
    ra.t = a->t;
    ra.y = a->y;
    ra.dy = a->dy;
    ra.obj_shift = obj_shift;
    ra.obj_gen = obj_gen;
    ra.precenter = plain_pre ? 1 : 0;
    const double dw = (a->max_freq - a->min_freq) / ((double)a->num_freqs * 1.0);
    if (split) {
        ra.rec_t = reinterpret_cast<double *>(ws + w.rec_t);
        ra.rec_a = reinterpret_cast<float4 *>(ws + w.rec_a);
        ra.rec_r = reinterpret_cast<double2 *>(ws + w.rec_r);
        ra.rec_n = reinterpret_cast<float *>(ws + w.rec_n);
        ra.step_rev = (double)T * dw * kInv2Pi;
        void (*fn)(RecArgs) = f64 ? (gen ? k_records_split<double, true> : k_records_split<double, false>)
                                  : (gen ? k_records_split<float, true> : k_records_split<float, false>);
        hipLaunchKernelGGL(fn, dim3(K), dim3(256), 0, st, ra);
        LS_HIP(hipGetLastError());
    } else if (gen) {
        ra.rec_y = ws + w.rec_y;
        ra.rec_w = ws + w.rec_w;
        if (f64)
            hipLaunchKernelGGL((k_records_plain<double, true>), dim3(K), dim3(256), 0, st, ra);
        else
            hipLaunchKernelGGL((k_records_plain<float, true>), dim3(K), dim3(256), 0, st, ra);
        LS_HIP(hipGetLastError());
    }

    // 3. the sweep
    SweepArgs sw{};
    sw.offsets = a->offsets;
    sw.order = a->order;
    sw.K = K;
    sw.nf = a->num_freqs;
    sw.chunks = w.chunks;
    sw.t = a->t;
    sw.y = plain_pre ? (const void *)(ws + w.rec_y) : a->y;
    sw.w = (!split && gen) ? (const void *)(ws + w.rec_w) : nullptr;
    if (f64) {
        const double fmin = a->min_freq, fmax = a->max_freq;
        sw.fmin_r = fmin;
        sw.fstep_r = (fmax - fmin) / ((double)a->num_freqs * 1.0);
    } else {
        const float fmin = (float)a->min_freq, fmax = (float)a->max_freq;
        sw.fmin_r = (double)fmin;
        sw.fstep_r = (double)(float)((double)(fmax - fmin) / ((double)a->num_freqs * 1.0));
    }
    sw.rec_t = split ? reinterpret_cast<const double *>(ws + w.rec_t) : nullptr;
    sw.rec_a = split ? reinterpret_cast<const float4 *>(ws + w.rec_a) : nullptr;
    sw.rec_r = split ? reinterpret_cast<const double2 *>(ws + w.rec_r) : nullptr;
    sw.rec_n = split ? reinterpret_cast<const float *>(ws + w.rec_n) : nullptr;
    sw.fmin_rev = a->min_freq * kInv2Pi;
    sw.df_rev = dw * kInv2Pi;
    sw.obj_gen = obj_gen;
    sw.obj_scale = obj_scale;
    sw.pgram = a->pgram;
    sw.part_pow = reinterpret_cast<double *>(ws + w.part_pow);
    sw.part_idx = reinterpret_cast<uint32_t *>(ws + w.part_idx);
    sw.top_k = (int)a->top_k;

    const uint64_t items = (uint64_t)K * w.chunks * w.slices;
    sw.total_items = items;
    sw.queue = nullptr;
    sw.slices = w.slices;
    sw.part_sums = w.slices > 1 ? reinterpret_cast<double *>(ws + w.part_sums) : nullptr;
    // items per workgroup fetch: the split forms reuse the staged records of
    // a small object across its consecutive chunks, when there are enough
    // items to keep every CU busy with a quarter as many workgroups
    sw.group = (split && items >= kGroupMinItems) ? kGroup : 1u;
    const uint64_t groups = (items + sw.group - 1) / sw.group;
    const bool persistent = a->schedule == LSAMD_SCHEDULE_PERSISTENT;
    uint32_t pgrid = 0;
    if (persistent) {
        sw.queue = reinterpret_cast<unsigned long long *>(ws + w.queue);
        int dev = 0, cus = 0, per_cu = 0;
        LS_HIP(hipGetDevice(&dev));
        LS_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        LS_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, reinterpret_cast<const void *>(te->fn), T, 0));
        pgrid = (uint32_t)std::max(1, cus * std::max(1, per_cu));
        if ((uint64_t)pgrid > groups) pgrid = (uint32_t)groups;
        hipLaunchKernelGGL(k_zero_counter, dim3(1), dim3(1), 0, st, sw.queue);
        LS_HIP(hipGetLastError());
    }
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    LS_HIP(hipStreamIsCapturing(st, &cap));
    const bool timing = g_timing && cap == hipStreamCaptureStatusNone;
    if (timing) {
        if (!g_ev[0]) {
            LS_HIP(hipEventCreate(&g_ev[0]));
            LS_HIP(hipEventCreate(&g_ev[1]));
        }
        LS_HIP(hipEventRecord(g_ev[0], st));
    }
    if (persistent) {
        sw.item_base = 0;
        void *kargs[] = {&sw};
        LS_HIP(hipLaunchKernel(reinterpret_cast<const void *>(te->fn), dim3(pgrid), dim3(T), kargs,
                               0, st));
    } else {
        // launches of whole groups, under 2^32 threads each
        const uint64_t max_groups = (0xFFFFFFFFull / (uint64_t)T) - 1;
        for (uint64_t g = 0; g < groups; g += max_groups) {
            const uint64_t cnt = std::min<uint64_t>(max_groups, groups - g);
            sw.item_base = g * sw.group;
            void *kargs[] = {&sw};
            LS_HIP(hipLaunchKernel(reinterpret_cast<const void *>(te->fn), dim3((uint32_t)cnt),
                                   dim3(T), kargs, 0, st));
        }
    }
    if (w.slices > 1) {
        if (!te->combine) return set_error(LSAMD_ERR_INVALID, "lsamd_sweep: no slice combiner");
        void *kargs[] = {&sw};
        LS_HIP(hipLaunchKernel(reinterpret_cast<const void *>(te->combine),
                               dim3((uint32_t)((uint64_t)K * w.chunks)), dim3(T), kargs, 0, st));
    }
    if (timing) {
        LS_HIP(hipEventRecord(g_ev[1], st));
        g_ev_pending = true;
    }
    (void)M;

    // 4. peaks
    if (a->top_k > 0) {
        ReduceArgs rd{};
        rd.K = K;
        rd.chunks = w.chunks;
        rd.top_k = a->top_k;
        rd.part_pow = sw.part_pow;
        rd.part_idx = sw.part_idx;
        rd.peak_index = a->peak_index;
        rd.peak_power = a->peak_power;
        rd.peak_period = a->peak_period;
        if (split || f64) {
            rd.pmin = a->min_freq;
            rd.pstep = dw;
        } else {
            const float fmin = (float)a->min_freq, fmax = (float)a->max_freq;
            rd.pmin = (double)fmin;
            rd.pstep = (double)(fmax - fmin) / ((double)a->num_freqs * 1.0);
        }
        if (f64)
            hipLaunchKernelGGL(k_reduce_peaks<double>, dim3(K), dim3(64), 0, st, rd);
        else
            hipLaunchKernelGGL(k_reduce_peaks<float>, dim3(K), dim3(64), 0, st, rd);
        LS_HIP(hipGetLastError());
    }
    return LSAMD_OK;
}

struct lsamd_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

int lsamd_graph_create(const lsamd_sweep_args *a, void *workspace, size_t workspace_bytes,
                       void *stream_v, lsamd_graph **out) {
    if (!out || !stream_v)
        return set_error(LSAMD_ERR_INVALID, "lsamd_graph_create: needs an output pointer and a "
                                            "non-default stream");
    *out = nullptr;
    hipStream_t st = static_cast<hipStream_t>(stream_v);
    LS_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    const int rc = lsamd_sweep(a, workspace, workspace_bytes, stream_v);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(st, &graph);
    lsamd_graph *g = new lsamd_graph();
    if (rc != LSAMD_OK || e != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        delete g;
        return rc != LSAMD_OK ? rc
                              : set_error(LSAMD_ERR_HIP, std::string("hipStreamEndCapture: ") +
                                                             hipGetErrorString(e));
    }
    g->graph = graph;
    const hipError_t ei = hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0);
    if (ei != hipSuccess) {
        (void)hipGraphDestroy(graph);
        delete g;
        return set_error(LSAMD_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
    }
    *out = g;
    return LSAMD_OK;
}

int lsamd_graph_launch(lsamd_graph *g, void *stream_v) {
    if (!g || !g->exec) return set_error(LSAMD_ERR_INVALID, "lsamd_graph_launch: null graph");
    hipStream_t st = static_cast<hipStream_t>(stream_v);
    if (g_timing) {
        if (!g_ev[0]) {
            LS_HIP(hipEventCreate(&g_ev[0]));
            LS_HIP(hipEventCreate(&g_ev[1]));
        }
        LS_HIP(hipEventRecord(g_ev[0], st));
    }
    LS_HIP(hipGraphLaunch(g->exec, st));
    if (g_timing) {
        LS_HIP(hipEventRecord(g_ev[1], st));
        g_ev_pending = true;
    }
    return LSAMD_OK;
}

void lsamd_graph_destroy(lsamd_graph *g) {
    if (!g) return;
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    delete g;
}

// The autotuner's trial sweeps: every object, at most 65536 bins, peaks only.
static lsamd_sweep_args autotune_sample(const lsamd_sweep_args *a) {
    lsamd_sweep_args s = *a;
    s.num_freqs = std::min<uint32_t>(a->num_freqs, 65536u);
    s.max_freq = a->min_freq + (a->max_freq - a->min_freq) * ((double)s.num_freqs / a->num_freqs);
    s.pgram = nullptr;
    s.top_k = 1;
    s.peak_index = nullptr;
    s.peak_power = nullptr;
    s.peak_period = nullptr;
    return s;
}

size_t lsamd_autotune_workspace_size(const lsamd_sweep_args *a) {
    if (!a || a->form < 0 || a->form >= kForms || a->variant < 0 || a->variant > 1) return 0;
    const lsamd_sweep_args s = autotune_sample(a);
    const int saved_T = g_tile_T[a->form][a->variant], saved_M = g_tile_M[a->form][a->variant];
    size_t need = 0;
    for (const TileEntry &te : kTiles) {
        if (te.form != a->form || te.variant != a->variant || te.dtype != a->dtype) continue;
        g_tile_T[a->form][a->variant] = te.T;
        g_tile_M[a->form][a->variant] = te.M;
        Workspace w;
        plan_workspace(&s, &w);
        need = std::max(need, w.bytes);
    }
    g_tile_T[a->form][a->variant] = saved_T;
    g_tile_M[a->form][a->variant] = saved_M;
    return need;
}

int lsamd_autotune(const lsamd_sweep_args *a, void *workspace, size_t workspace_bytes,
                   void *stream_v, int *threads_out, int *fpt_out) {
    if (!a) return set_error(LSAMD_ERR_INVALID, "lsamd_autotune: null args");
    if (a->form < 0 || a->form >= kForms || a->variant < 0 || a->variant > 1)
        return set_error(LSAMD_ERR_INVALID, "lsamd_autotune: bad form/variant");
    hipStream_t st = static_cast<hipStream_t>(stream_v);
    lsamd_sweep_args s = autotune_sample(a);
    const int form = a->form, variant = a->variant;
    const int saved_T = g_tile_T[form][variant], saved_M = g_tile_M[form][variant];
    const bool saved_timing = g_timing;
    hipEvent_t e0, e1;
    LS_HIP(hipEventCreate(&e0));
    LS_HIP(hipEventCreate(&e1));
    float best_ms = 1e30f;
    int best_T = 0, best_M = 0;
    for (const TileEntry &te : kTiles) {
        if (te.form != form || te.variant != variant || te.dtype != a->dtype) continue;
        g_tile_T[form][variant] = te.T;
        g_tile_M[form][variant] = te.M;
        Workspace w;
        plan_workspace(&s, &w);
        if (w.bytes > workspace_bytes) continue;
        g_timing = false;
        int rc = lsamd_sweep(&s, workspace, workspace_bytes, stream_v);  // warm-up
        if (rc) break;
        LS_HIP(hipEventRecord(e0, st));
        rc = lsamd_sweep(&s, workspace, workspace_bytes, stream_v);
        if (rc) break;
        LS_HIP(hipEventRecord(e1, st));
        LS_HIP(hipEventSynchronize(e1));
        float ms = 0.0f;
        LS_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best_ms) {
            best_ms = ms;
            best_T = te.T;
            best_M = te.M;
        }
    }
    g_timing = saved_timing;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (best_T == 0) {
        g_tile_T[form][variant] = saved_T;
        g_tile_M[form][variant] = saved_M;
        return lsamd_last_error_code() ? lsamd_last_error_code()
                                       : set_error(LSAMD_ERR_INVALID, "lsamd_autotune: no tile fits "
                                                                      "the workspace");
    }
    g_tile_T[form][variant] = best_T;
    g_tile_M[form][variant] = best_M;
    if (threads_out) *threads_out = best_T;
    if (fpt_out) *fpt_out = best_M;
    return LSAMD_OK;
}

}  // extern "C"
