// The stream-ordered device entry point lsamd_sweep() of the MI355X-native
// Lomb-Scargle sweep: per-object statistics, per-observation records, the
// sweep kernel of the form and tile (ls_sweep_plain.hip, ls_sweep_split.hip),
// the per-object peak reduction, plus graph capture and the tile autotuner.
//
// Work decomposition: work items (object, frequency chunk), one workgroup
// each, or taken in groups of consecutive chunks.  A chunk is T*M
// consecutive frequency bins; in the VALU kernels thread tid owns bins
// chunk*T*M + m*T + tid (m < M).  The object's observations are staged
// through LDS T at a time and read back as wave-wide broadcasts; all sums
// live in registers.  The matrix-core tile (64 x 16) maps a chunk to a
// 32 x 32 block of bins instead (see k_sweep_mfma).
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
//   split32 - the same seed in FP32; the VALU tiles step across the bins in
//             FP32 (rotation, recurrence) with FP32 partial sums over 32-64
//             observations folded into FP64; the default tiles form the
//             sums on the FP16 matrix cores as row x column factor products
//             (head/tail split), folded every 256 observations (standard:
//             FP32 fold; generalized: the y sums Zy into FP64, the
//             weight-only Zw and Z2 into FP32).
// Both split forms use the tau-free FP64 power
//     standard:    P = (SS XC^2 - 2 CS XC XS + CC XS^2) / (2 (CC SS - CS^2))
//     generalized: P = (SS YC^2 - 2 CS YC YS + CC YS^2) / ((CC SS - CS^2) YY)
// (centred weighted sums), equal to the reference's tau form because the
// 2x2 least-squares projection is rotation invariant.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <string>

#include "ls_device.h"

namespace lsamd {

namespace {

// ---------------------------------------------------------------------------
// Kernel arguments of the small kernels around the sweep
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
        double Wf = 0.0, Ys = 0.0, YYs = 0.0, Wd = 0.0, Yd = 0.0, YYd = 0.0;
        for (uint64_t i = b; i < e; ++i) {
            const double qi = sqrt(1.0 / ((double)dy[i] * (double)dy[i]) / W);
            const double yc = (double)y[i] - ybar;
            const double q = (double)(float)qi;
            const float bq = (float)(qi * yc);
            Wf += q * q;
            Ys += (double)bq * q;
            YYs += q * q * yc * yc;
            const double wn = 1.0 / ((double)dy[i] * (double)dy[i]) / W;
            Wd += wn;
            Yd += wn * yc;
            YYd += wn * yc * yc;
        }
        const double Yb = Ys / Wf;
        a.obj_gen[kGen * k + 0] = Wf;
        a.obj_gen[kGen * k + 1] = Yb;
        a.obj_gen[kGen * k + 2] = YYs / Wf - Yb * Yb;
        a.obj_gen[kGen * k + 3] = ybar;
        a.obj_gen[kGen * k + 4] = W;
        a.obj_gen[kGen * k + 5] = Wd;
        a.obj_gen[kGen * k + 6] = Yd / Wd;
        a.obj_gen[kGen * k + 7] = YYd / Wd - (Yd / Wd) * (Yd / Wd);
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
        double Wf = 0.0, Ys = 0.0, YYs = 0.0, Wd = 0.0, Yd = 0.0, YYd = 0.0;
        for (uint64_t i = b + lane; i < e; i += 64) {
            const double qi = sqrt(1.0 / ((double)dy[i] * (double)dy[i]) / W);
            const double yc = (double)y[i] - ybar;
            const double q = (double)(float)qi;
            const float bq = (float)(qi * yc);
            Wf += q * q;
            Ys += (double)bq * q;
            YYs += q * q * yc * yc;
            const double wn = 1.0 / ((double)dy[i] * (double)dy[i]) / W;
            Wd += wn;
            Yd += wn * yc;
            YYd += wn * yc * yc;
        }
        Wf = wave_sum64(Wf);
        Ys = wave_sum64(Ys);
        YYs = wave_sum64(YYs);
        Wd = wave_sum64(Wd);
        Yd = wave_sum64(Yd);
        YYd = wave_sum64(YYd);
        if (lane == 0) {
            const double Yb = Ys / Wf;
            a.obj_gen[kGen * k + 0] = Wf;
            a.obj_gen[kGen * k + 1] = Yb;
            a.obj_gen[kGen * k + 2] = YYs / Wf - Yb * Yb;
            a.obj_gen[kGen * k + 3] = ybar;
            a.obj_gen[kGen * k + 4] = W;
            a.obj_gen[kGen * k + 5] = Wd;
            a.obj_gen[kGen * k + 6] = Yd / Wd;
            a.obj_gen[kGen * k + 7] = YYd / Wd - (Yd / Wd) * (Yd / Wd);
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

__global__ void k_zero_counters(unsigned long long *c, uint32_t n) {
    if (threadIdx.x < n) c[threadIdx.x] = 0ull;
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
// Tile registry: the plain and split kernel tables
// ---------------------------------------------------------------------------
// The compiled kernels of one form.
const TileEntry *tiles_of(int form, int *count) {
    *count = form == LSAMD_FORM_PLAIN ? kPlainTileCount : kSplitTileCount;
    return form == LSAMD_FORM_PLAIN ? kPlainTiles : kSplitTiles;
}

const TileEntry *find_tile(int form, int variant, int dtype, int T, int M) {
    int n = 0;
    const TileEntry *tab = tiles_of(form, &n);
    for (int i = 0; i < n; ++i) {
        const TileEntry &e = tab[i];
        if (e.form == form && e.variant == variant && e.dtype == dtype && e.T == T && e.M == M)
            return &e;
    }
    return nullptr;
}

// Process-wide default tile per [form][variant], packed as T << 16 | M
// (0 = built-in), for sweeps whose args leave the tile at 0.  Sweeps read it
// once when they are planned; per-call tiles never touch it.
constexpr int kForms = 3;
std::atomic<uint32_t> g_default_tile[kForms][2];

// Built-in shapes: the fastest of the tile sweep on the 1000-object x 1e6
// batches (MI355X): split 256x16 (standard) / 256x8 (generalized); split32
// 64x16, the matrix-core tiles (the fastest VALU tiles: 256x8 standard,
// 256x4 generalized).  lsamd_autotune re-picks per workload.
Tile default_tile(int form, int variant) {
    if (form == LSAMD_FORM_SPLIT) return variant ? Tile{256, 8} : Tile{256, 16};
    if (form == LSAMD_FORM_SPLIT32) return Tile{64, 16};
    return Tile{256, 1};
}

// Kernel timing, per host thread: the events bracketing this thread's last
// timed sweep (created on the device of that sweep's stream).
struct Timing {
    bool enabled = false;
    double last_ms = -1.0;
    int device = -1;
    hipEvent_t ev[2] = {nullptr, nullptr};
    bool pending = false;
};
thread_local Timing t_timing;

// Events of the calling thread on the current device, created on first use.
hipError_t timing_events(Timing &tm) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (tm.ev[0] && tm.device == dev) return hipSuccess;
    for (hipEvent_t &ev : tm.ev) {
        if (ev) (void)hipEventDestroy(ev);
        ev = nullptr;
    }
    tm.pending = false;
    for (hipEvent_t &ev : tm.ev)
        if ((e = hipEventCreate(&ev)) != hipSuccess) return e;
    tm.device = dev;
    return hipSuccess;
}

// Events owned by one scope (the autotuner's trial timings).
struct ScopedEvents {
    hipEvent_t ev[2] = {nullptr, nullptr};
    ~ScopedEvents() {
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
    }
};

size_t align_up(size_t x) { return (x + 255) & ~(size_t)255; }

}  // namespace

Tile current_tile(int form, int variant, int dtype) {
    (void)dtype;
    if (form < 0 || form >= kForms || variant < 0 || variant > 1) return Tile{0, 0};
    const uint32_t v = g_default_tile[form][variant].load(std::memory_order_relaxed);
    if (v) return Tile{(int)(v >> 16), (int)(v & 0xFFFFu)};
    return default_tile(form, variant);
}

Tile tile_of(const lsamd_sweep_args *a) {
    if (a->tile_threads > 0 && a->tile_fpt > 0) return Tile{a->tile_threads, a->tile_fpt};
    return current_tile(a->form, a->variant, a->dtype);
}

bool tile_supported(int form, int variant, int threads, int fpt) {
    return find_tile(form, variant, LSAMD_F32, threads, fpt) != nullptr &&
           find_tile(form, variant, LSAMD_F64, threads, fpt) != nullptr;
}

int plan_workspace(const lsamd_sweep_args *a, Workspace *w) {
    *w = Workspace{};
    const Tile tile = tile_of(a);
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
    const TileEntry *te = find_tile(a->form, a->variant, a->dtype, tile.threads, tile.fpt);
    if (te && te->prep) {  // matrix-core tile: y scales and column-factor fragments
        w->obj_ys = take(8 * K);
        const size_t frags = a->variant == LSAMD_GENERALIZED ? kColfacFragsGen : kColfacFragsStd;
        w->colfac = take((size_t)16 * 64 * frags * (n / 16 + K + 1));
        w->rowfac = take((size_t)16 * 64 * kRowfacFrags * (n / 16 + K + 1));
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
    w->queue = take(8 * kQueues);
    w->bytes = off + 256;
    return LSAMD_OK;
}

}  // namespace lsamd

using namespace lsamd;

// Workspace of lsamd_sweep_static: a device array of LSAMD_STATIC_SCRATCH_BYTES
// on every GPU the library runs on (zero-filled at load, held until the
// process ends).  Host code uses only its per-device address from
// hipGetSymbolAddress.  Every sweep writes what it reads (the persistent
// schedule's counter is cleared by a kernel), so no call depends on the
// previous call's contents.
namespace {
constexpr size_t kStaticScratchBytes = LSAMD_STATIC_SCRATCH_BYTES;
constexpr int kScratchDevices = 64;
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];
std::atomic<void *> g_scratch_ptr[kScratchDevices];
}  // namespace

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
        g_default_tile[form][variant].store(0u);
        return LSAMD_OK;
    }
    if (!tile_supported(form, variant, threads, fpt))
        return set_error(LSAMD_ERR_INVALID, "lsamd_set_tile: tile shape " + std::to_string(threads) +
                                                "x" + std::to_string(fpt) + " not compiled");
    g_default_tile[form][variant].store((uint32_t)threads << 16 | (uint32_t)fpt);
    return LSAMD_OK;
}

void lsamd_get_tile(int form, int variant, int dtype, int *threads, int *fpt) {
    const Tile t = current_tile(form, variant, dtype);
    if (threads) *threads = t.threads;
    if (fpt) *fpt = t.fpt;
}

void lsamd_set_timing(int enabled) { t_timing.enabled = enabled != 0; }

double lsamd_last_sweep_kernel_ms(void) {
    Timing &tm = t_timing;
    if (tm.pending) {
        float ms = -1.0f;
        if (hipEventElapsedTime(&ms, tm.ev[0], tm.ev[1]) == hipSuccess) tm.last_ms = ms;
        tm.pending = false;
    }
    return tm.last_ms;
}

#define LS_HIP(call)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess)                                                                \
            return set_error(LSAMD_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

// The sweep into a given workspace; from_scratch: the workspace is the static
// device array of lsamd_sweep_static (no kernel timing, which would create
// events).
static int sweep_into(const lsamd_sweep_args *a, void *workspace, size_t workspace_bytes, void *stream_v,
                      bool from_scratch) {
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
    if ((a->tile_threads != 0 || a->tile_fpt != 0) &&
        !find_tile(a->form, a->variant, a->dtype, a->tile_threads, a->tile_fpt))
        return set_error(LSAMD_ERR_INVALID, "lsamd_sweep: tile shape " + std::to_string(a->tile_threads) +
                                                "x" + std::to_string(a->tile_fpt) + " not compiled");
    Workspace w;
    plan_workspace(a, &w);
    if (from_scratch && w.bytes > workspace_bytes)
        return set_error(LSAMD_ERR_INVALID, "lsamd_sweep_static: these arguments need " +
                                                std::to_string(w.bytes) + " bytes of workspace, more than the " +
                                                std::to_string(workspace_bytes) +
                                                "-byte static scratch; use lsamd_sweep with a workspace");
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
    sw.dy = gen ? a->dy : nullptr;
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
    sw.colfac = w.colfac ? reinterpret_cast<uint4 *>(ws + w.colfac) : nullptr;
    sw.rowfac = w.rowfac ? reinterpret_cast<uint4 *>(ws + w.rowfac) : nullptr;
    sw.obj_ys = w.obj_ys ? reinterpret_cast<float *>(ws + w.obj_ys) : nullptr;

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
    // about one object's groups per XCD run (grid schedule and persistent queues)
    sw.xcd_run = (uint32_t)std::max<uint64_t>(1, (uint64_t)w.chunks * w.slices / sw.group);
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
        hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(64), 0, st, sw.queue, kQueues);
        LS_HIP(hipGetLastError());
    }
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    LS_HIP(hipStreamIsCapturing(st, &cap));
    Timing &tm = t_timing;
    const bool timing = tm.enabled && cap == hipStreamCaptureStatusNone && !from_scratch;
    if (timing) {
        LS_HIP(timing_events(tm));
        LS_HIP(hipEventRecord(tm.ev[0], st));
    }
    if (te->prep) {  // per-object set-up of the tile, inside the timed sweep
        // wavefronts per object: enough for the mean object's 16-observation
        // blocks, up to 256 (one object of many observations)
        const uint64_t blocks = (a->num_obs / 16 + K - 1) / K + 1;
        const uint32_t gy = (uint32_t)std::min<uint64_t>(256, blocks);
        void *kargs[] = {&sw};
        LS_HIP(hipLaunchKernel(reinterpret_cast<const void *>(te->prep), dim3(K, gy), dim3(64), kargs,
                               0, st));
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
        LS_HIP(hipEventRecord(tm.ev[1], st));
        tm.pending = true;
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

int lsamd_sweep(const lsamd_sweep_args *a, void *workspace, size_t workspace_bytes, void *stream_v) {
    return sweep_into(a, workspace, workspace_bytes, stream_v, false);
}

int lsamd_sweep_static(const lsamd_sweep_args *a, void *stream_v) {
    int dev = 0;
    LS_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kScratchDevices)
        return set_error(LSAMD_ERR_INVALID, "lsamd_sweep_static: device " + std::to_string(dev) +
                                                " beyond the " + std::to_string(kScratchDevices) +
                                                " devices it keeps a scratch address for");
    void *ptr = g_scratch_ptr[dev].load(std::memory_order_acquire);
    if (!ptr) {
        // each device holds its own copy of the array: its address on the
        // current device (the host-side name is not a device address)
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return set_error(LSAMD_ERR_HIP, std::string("lsamd_sweep_static: hipGetSymbolAddress(scratch): ") +
                                                hipGetErrorString(e));
        g_scratch_ptr[dev].store(ptr, std::memory_order_release);
    }
    return sweep_into(a, ptr, kStaticScratchBytes, stream_v, true);
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
    Timing &tm = t_timing;
    if (tm.enabled) {
        LS_HIP(timing_events(tm));
        LS_HIP(hipEventRecord(tm.ev[0], st));
    }
    LS_HIP(hipGraphLaunch(g->exec, st));
    if (tm.enabled) {
        LS_HIP(hipEventRecord(tm.ev[1], st));
        tm.pending = true;
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
    lsamd_sweep_args s = autotune_sample(a);
    size_t need = 0;
    int ntab = 0;
    const TileEntry *tab = tiles_of(a->form, &ntab);
    for (int i = 0; i < ntab; ++i) {
        const TileEntry &te = tab[i];
        if (te.form != a->form || te.variant != a->variant || te.dtype != a->dtype) continue;
        s.tile_threads = te.T;
        s.tile_fpt = te.M;
        Workspace w;
        plan_workspace(&s, &w);
        need = std::max(need, w.bytes);
    }
    return need;
}

// Tiles lsamd_autotune may pick: all compiled shapes, except the VALU tiles of
// generalized split32 (FP32 rotation and sums), whose worst power error on the
// 1000-object generalized batch with 2000-day spans is 1.07e-5, above the
// form's 1e-5 bound (the matrix-core tile: 8.4e-6).  They stay available by
// explicit choice (tile_threads / tile_fpt, lsamd_set_tile).
static bool autotune_candidate(const TileEntry &te) {
    return !(te.form == LSAMD_FORM_SPLIT32 && te.variant == LSAMD_GENERALIZED && te.T != 64);
}

int lsamd_autotune(const lsamd_sweep_args *a, void *workspace, size_t workspace_bytes,
                   void *stream_v, int *threads_out, int *fpt_out) {
    if (!a) return set_error(LSAMD_ERR_INVALID, "lsamd_autotune: null args");
    if (a->form < 0 || a->form >= kForms || a->variant < 0 || a->variant > 1)
        return set_error(LSAMD_ERR_INVALID, "lsamd_autotune: bad form/variant");
    hipStream_t st = static_cast<hipStream_t>(stream_v);
    lsamd_sweep_args s = autotune_sample(a);
    // the trial sweeps are timed here, not by the caller's thread timing
    Timing &tm = t_timing;
    const bool saved_timing = tm.enabled;
    tm.enabled = false;
    struct Restore {
        Timing &tm;
        bool on;
        ~Restore() { tm.enabled = on; }
    } restore{tm, saved_timing};
    ScopedEvents ev;
    LS_HIP(hipEventCreate(&ev.ev[0]));
    LS_HIP(hipEventCreate(&ev.ev[1]));
    float best_ms = 1e30f;
    int best_T = 0, best_M = 0;
    int ntab = 0;
    const TileEntry *tab = tiles_of(a->form, &ntab);
    for (int i = 0; i < ntab; ++i) {
        const TileEntry &te = tab[i];
        if (te.form != a->form || te.variant != a->variant || te.dtype != a->dtype) continue;
        if (!autotune_candidate(te)) continue;
        s.tile_threads = te.T;
        s.tile_fpt = te.M;
        Workspace w;
        plan_workspace(&s, &w);
        if (w.bytes > workspace_bytes) continue;
        int rc = lsamd_sweep(&s, workspace, workspace_bytes, stream_v);  // warm-up
        if (rc) return rc;
        LS_HIP(hipEventRecord(ev.ev[0], st));
        rc = lsamd_sweep(&s, workspace, workspace_bytes, stream_v);
        if (rc) return rc;
        LS_HIP(hipEventRecord(ev.ev[1], st));
        LS_HIP(hipEventSynchronize(ev.ev[1]));
        float ms = 0.0f;
        LS_HIP(hipEventElapsedTime(&ms, ev.ev[0], ev.ev[1]));
        if (ms < best_ms) {
            best_ms = ms;
            best_T = te.T;
            best_M = te.M;
        }
    }
    if (best_T == 0)
        return set_error(LSAMD_ERR_INVALID, "lsamd_autotune: no tile fits the workspace");
    if (threads_out) *threads_out = best_T;
    if (fpt_out) *fpt_out = best_M;
    return LSAMD_OK;
}

}  // extern "C"
