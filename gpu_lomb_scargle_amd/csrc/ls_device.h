// Device-side pieces shared by the sweep kernels (ls_sweep_plain.hip,
// ls_sweep_split.hip) and the host side of the sweep (ls_sweep.hip): the
// sweep's launch arguments, the work-item schedule, per-chunk peak lists and
// the tile registry entry.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "ls_internal.h"

namespace lsamd {

struct SweepArgs {
    const uint64_t *offsets;
    const uint32_t *order;
    uint32_t K, nf, chunks;
    uint64_t item_base;        // first item of this launch
    uint64_t total_items;      // K * chunks
    uint32_t group;            // consecutive items per workgroup fetch (a power of two)
    uint32_t xcd_run;          // grid schedule: groups kept together on one XCD (0: blockIdx order)
    uint32_t slices;           // split forms: observation slices per (object, chunk), >= 1
    double *part_sums;         // slices > 1: per-slice FP64 sums [K*chunks][slices][NS][M][T]
    unsigned long long *queue; // persistent schedule: kQueues group counters (one per XCD), else null
    // plain (t, y, w); split forms: the caller's y and dy (DTYPE), read by
    // the FP64 recomputation of ill-conditioned bins and small objects
    const void *t, *y, *w, *dy;
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
    // matrix-core tile: per-object column- and row-factor fragments, y scale
    uint4 *colfac;
    uint4 *rowfac;
    float *obj_ys;  // per object: y scale, weight scale
};

typedef void (*SweepFn)(SweepArgs);

// One compiled sweep kernel: form, variant, output type and tile shape.
struct TileEntry {
    int form, variant, dtype, T, M;
    SweepFn fn;
    SweepFn combine;  // split forms: sums the observation slices (null: plain)
    SweepFn prep;     // per-object set-up launched before the sweep (K x G workgroups of 64), or null
};
extern const TileEntry kPlainTiles[];  // ls_sweep_plain.hip
extern const int kPlainTileCount;
extern const TileEntry kSplitTiles[];  // ls_sweep_split.hip
extern const int kSplitTileCount;

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
// Matrix-core tile: column-factor fragments per 16 observations (standard:
// y v and w; generalized: (w/W) v, (w/W) y_c v and (w/W) w; real and
// imaginary head and tail of each).
constexpr int kColfacFragsStd = 8, kColfacFragsGen = 12;
// Matrix-core tile: row-step phasors e^{2 pi i 32 a df t} per 16 observations
// (lane a + 32 h: observations 8 h .. 8 h + 7 as 4 x two complex FP32).
constexpr int kRowfacFrags = 4;
constexpr int kPeaks = 3;      // peaks tracked per object (top-1 and top-3)
// doubles per object of split-generalized constants: [0..2] Wf, Yb, YY over
// the FP32-rounded weights the VALU split sweeps apply; [3] ybar, [4] W (sum
// of 1/dy^2); [5..7] Wf, Yb, YY over the FP64 weights w / W (matrix-core tile)
constexpr int kGen = 8;
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
// nothing for top_k = 0.  Each thread adds its bins in ascending order, so an
// equal power never displaces an earlier bin and a strict comparison of the
// powers decides.
// Out-of-order insertion (bins that are not the lane's next in index order):
// the index breaks ties, as the in-order strict comparison of peaks_add does.
__device__ __forceinline__ void peaks_merge(Peaks &L, int top_k, double p, uint32_t i) {
    if (top_k == 1) {
        if (better(p, i, L.p[0], L.i[0])) {
            L.p[0] = p;
            L.i[0] = i;
        }
    } else if (top_k) {
        peaks_insert(L, p, i);
    }
}

__device__ __forceinline__ void peaks_add(Peaks &L, int top_k, double p, uint32_t i) {
    if (top_k == 1) {
        if (p > L.p[0]) {
            L.p[0] = p;
            L.i[0] = i;
        }
    } else if (top_k) {
        const bool c0 = p > L.p[0], c1 = p > L.p[1], c2 = p > L.p[2];
        const double p2 = c1 ? L.p[1] : (c2 ? p : L.p[2]);
        const uint32_t i2 = c1 ? L.i[1] : (c2 ? i : L.i[2]);
        const double p1 = c0 ? L.p[0] : (c1 ? p : L.p[1]);
        const uint32_t i1 = c0 ? L.i[0] : (c1 ? i : L.i[1]);
        L.p[0] = c0 ? p : L.p[0];
        L.i[0] = c0 ? i : L.i[0];
        L.p[1] = p1;
        L.i[1] = i1;
        L.p[2] = p2;
        L.i[2] = i2;
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

// Best (power, bin) of a whole workgroup, valid in every thread; `slot`
// selects the LDS entries used, so that consecutive calls need no barrier
// between them.
template <int T>
__device__ __forceinline__ void block_best_all(double &p, uint32_t &i, double (*s_p)[kPeaks],
                                               uint32_t (*s_i)[kPeaks], int slot) {
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
            s_p[wave][slot] = p;
            s_i[wave][slot] = i;
        }
        __syncthreads();
        p = s_p[0][slot];
        i = s_i[0][slot];
#pragma unroll
        for (int w = 1; w < T / 64; ++w)
            if (better(s_p[w][slot], s_i[w][slot], p, i)) {
                p = s_p[w][slot];
                i = s_i[w][slot];
            }
    }
}

// Top three (power, bin) of a whole workgroup from every thread's own
// sorted top three: three rounds of "best head, then its owner drops it",
// i.e. three scans that exclude the bins already taken.  The result is
// returned in every thread.
template <int T>
__device__ __forceinline__ void block_top3(Peaks &L, double (*s_p)[kPeaks],
                                           uint32_t (*s_i)[kPeaks], Peaks &W) {
#pragma unroll
    for (int r = 0; r < kPeaks; ++r) {
        double p = L.p[0];
        uint32_t i = L.i[0];
        block_best_all<T>(p, i, s_p, s_i, r);
        W.p[r] = p;
        W.i[r] = i;
        if (L.p[0] == p && L.i[0] == i) {  // the owner (or an empty (0, 0) head)
            L.p[0] = L.p[1];
            L.i[0] = L.i[1];
            L.p[1] = L.p[2];
            L.i[1] = L.i[2];
            L.p[2] = 0.0;
            L.i[2] = 0u;
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
// (item_base + grid_group * group, below), or, with queues, persistent
// workgroups that take groups until the items run out.
//
// Persistent schedule: kQueues counters, one per XCD.  Groups form runs of
// L = xcd_run (about one object's chunks); queue q holds runs q, 15 - q,
// 16 + q, 31 - q, ... (the grid schedule's XCD assignment), so its i-th group
// is run 8 rho + (rho odd ? 7 - q : q), rho = i / L, offset i % L: every group
// belongs to exactly one queue, and a queue's groups ascend with i.  A
// workgroup starts on the queue of its own XCD (blockIdx.x % 8, the dispatch
// order) and, once that queue is exhausted, takes groups from the next ones in
// turn, so that an object's tables stay in one XCD's L2 while its work lasts
// and no XCD idles while another has work.
constexpr uint32_t kQueues = 8;
__device__ __forceinline__ uint64_t queue_group(const SweepArgs &a, uint32_t q, uint64_t i) {
    const uint64_t L = a.xcd_run ? a.xcd_run : 1u;
    const uint64_t rho = i / L, x = (rho & 1u) ? 7u - q : q;
    return (8u * rho + x) * L + (i - rho * L);
}
__device__ __forceinline__ uint64_t fetch_item(const SweepArgs &a, uint64_t *s_item, bool first) {
    __shared__ uint32_t s_queue, s_tried;  // this workgroup's current queue, queues found empty
    __syncthreads();  // every thread is done with the previous item and its s_item read
    if (threadIdx.x == 0) {
        if (first) {
            s_queue = blockIdx.x & (kQueues - 1u);
            s_tried = 0u;
        }
        const uint64_t groups = (a.total_items - a.item_base + a.group - 1u) / a.group;
        uint64_t item = ~0ull;
        while (s_tried < kQueues) {
            const uint64_t i = atomicAdd(&a.queue[s_queue], 1ull);
            const uint64_t g = queue_group(a, s_queue, i);
            if (g < groups) {
                item = a.item_base + g * a.group;
                break;
            }
            s_queue = (s_queue + 1u) & (kQueues - 1u);  // exhausted for good: groups ascend with i
            ++s_tried;
        }
        *s_item = item;
    }
    __syncthreads();
    return *s_item;
}
// Grid schedule: the 8 XCDs receive workgroups round-robin (workgroup w goes
// to XCD w % 8), each with its own L2.  Runs of L = xcd_run consecutive
// groups (about one object's chunks) go to one XCD, and consecutive runs to
// consecutive XCDs, in alternating direction per round of 8 runs (objects
// come longest first): workgroup w = 8 i + x takes run 8 r + x' (r = i / L,
// x' = x, or 7 - x for odd r), group i % L of it; a bijection on the whole
// rounds of 8 L groups (the rest in blockIdx order).  An object's tables are
// then fetched into one L2, and the XCDs get equal shares of the work.
__device__ __forceinline__ uint64_t grid_group(const SweepArgs &a) {
    const uint64_t w = blockIdx.x, L = a.xcd_run;
    if (L == 0) return w;
    const uint64_t round = 8u * L, full = ((uint64_t)gridDim.x / round) * round;
    if (w >= full) return w;
    const uint64_t i = w >> 3, r = i / L, x = (r & 1u) ? 7u - (w & 7u) : (w & 7u);
    return r * round + x * L + i % L;
}
__device__ __forceinline__ uint64_t first_item(const SweepArgs &a, uint64_t *s_item) {
    return a.queue ? fetch_item(a, s_item, true) : a.item_base + grid_group(a) * a.group;
}
__device__ __forceinline__ uint64_t next_item(const SweepArgs &a, uint64_t *s_item, uint64_t item) {
    const uint64_t n = item + 1;
    if (((n - a.item_base) & (uint64_t)(a.group - 1)) != 0) return n;  // same group
    return a.queue ? fetch_item(a, s_item, false) : ~0ull;
}

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
    Peaks W;
    block_top3<T>(L, s_p, s_i, W);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int r = 0; r < kPeaks; ++r) {
            a.part_pow[o + r] = W.p[r];
            a.part_idx[o + r] = W.i[r];
        }
    }
}

}  // namespace

}  // namespace lsamd
