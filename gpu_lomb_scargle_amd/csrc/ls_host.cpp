// Host orchestration of the MI355X-native Lomb-Scargle library: error state,
// segmentation of the observation stream into objects, reference-faithful
// pre-centring, device buffers, memory-budget batching and the pinned,
// double-buffered copy of the periodogram back into the caller's memory.
//
// Replaces the reference's GPU host path: computeObjectRanges
// (source/main.cu:468-510), computeNumBatches (main.cu:423-466),
// batchGPULSWrapper / batchGPULS (main.cu:742-1280), GPULSOneObject
// (main.cu:570-720) and pinnedMemoryCopyDtoH (main.cu:513-559).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "ls_internal.h"

namespace lsamd {

// ---------------------------------------------------------------------------
// error state
// ---------------------------------------------------------------------------
namespace {
thread_local int t_err_code = LSAMD_OK;
thread_local std::string t_err_msg;
}  // namespace

int set_error(int code, const std::string &msg) {
    t_err_code = code;
    t_err_msg = msg;
    return code;
}

void clear_error() {
    t_err_code = LSAMD_OK;
    t_err_msg.clear();
}

// ---------------------------------------------------------------------------
// segmentation: objects are runs of equal ids (main.cu:468-510)
// ---------------------------------------------------------------------------
uint32_t count_runs(const uint32_t *id, uint64_t n) {
    if (n == 0) return 0;
    uint32_t k = 1;
    for (uint64_t i = 1; i < n; ++i) k += id[i] != id[i - 1];
    return k;
}

uint32_t run_offsets(const uint32_t *id, uint64_t n, uint64_t *off, uint32_t *ids) {
    if (n == 0) return 0;
    uint32_t k = 0;
    off[0] = 0;
    for (uint64_t i = 1; i < n; ++i) {
        if (id[i] != id[i - 1]) {
            if (ids) ids[k] = id[i - 1];
            off[++k] = i;
        }
    }
    if (ids) ids[k] = id[n - 1];
    off[++k] = n;
    return k;
}

// updateYerrorfactor (main.cu:1631-1662) with the reductions summed in index
// order; w = 1/sqrt(dy) is the reference's weight for this step.
template <typename R>
void precenter_host(R *y, const R *dy, uint64_t n) {
    R sumw = 0, dot = 0;
    for (uint64_t i = 0; i < n; ++i) sumw += (R)(1.0 / (double)std::sqrt(dy[i]));
    for (uint64_t i = 0; i < n; ++i) dot += (R)(1.0 / (double)std::sqrt(dy[i])) * y[i];
    const R shift = dot / sumw;
    for (uint64_t i = 0; i < n; ++i) y[i] = y[i] - shift;
}
template void precenter_host<float>(float *, const float *, uint64_t);
template void precenter_host<double>(double *, const double *, uint64_t);

// ---------------------------------------------------------------------------
// device contexts: per device, one stream and grow-only buffers, reused
// across calls
// ---------------------------------------------------------------------------
namespace {

struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
};

struct Ctx {
    std::mutex mu;
    bool ready = false;
    int device = -1;
    hipStream_t stream = nullptr;       // uploads and sweeps
    hipStream_t copy_stream = nullptr;  // periodogram copy-back
    DevBuf t, y, dy, off, order, ws, pgram, pidx, ppow, pper;
    void *pinned[2] = {nullptr, nullptr};
    size_t pinned_cap = 0;
    hipEvent_t ev[2] = {nullptr, nullptr};     // staging chunks (copy stream)
    hipEvent_t slice[2] = {nullptr, nullptr};  // sweep of a slice done (sweep stream)
};

// One context per device, for the device current on the calling thread.
// Calls on one device serialise on its context; calls on different devices
// (one host thread each) run side by side.  Contexts are never destroyed:
// HIP may be torn down first at exit.
Ctx *ctx_for_current_device(int *err) {
    static std::mutex mu;
    static std::map<int, Ctx *> *table = new std::map<int, Ctx *>();
    int dev = 0;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        *err = set_error(LSAMD_ERR_NODEVICE, "no HIP device visible");
        return nullptr;
    }
    const hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        *err = set_error(LSAMD_ERR_HIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(mu);
    Ctx *&c = (*table)[dev];
    if (!c) {
        c = new Ctx();
        c->device = dev;
    }
    return c;
}

#define HIP_OK(call)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return set_error(e_ == hipErrorOutOfMemory ? LSAMD_ERR_NOMEM : LSAMD_ERR_HIP,     \
                             std::string(#call) + ": " + hipGetErrorString(e_));              \
    } while (0)

int ctx_init(Ctx &c) {
    if (c.ready) return LSAMD_OK;
    HIP_OK(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
    HIP_OK(hipStreamCreateWithFlags(&c.copy_stream, hipStreamNonBlocking));
    for (int i = 0; i < 2; ++i) {
        HIP_OK(hipEventCreateWithFlags(&c.ev[i], hipEventDisableTiming));
        HIP_OK(hipEventCreateWithFlags(&c.slice[i], hipEventDisableTiming));
    }
    c.ready = true;
    return LSAMD_OK;
}

int ensure(DevBuf &b, size_t bytes) {
    if (bytes == 0) bytes = 256;
    if (b.cap >= bytes) return LSAMD_OK;
    if (b.p) {
        HIP_OK(hipFree(b.p));
        b.p = nullptr;
        b.cap = 0;
    }
    HIP_OK(hipMalloc(&b.p, bytes));
    b.cap = bytes;
    return LSAMD_OK;
}

constexpr size_t kPinnedChunk = size_t(64) << 20;  // 64 MiB per staging buffer

// Device -> pageable host copy on the copy stream through two pinned staging
// buffers, so the DMA of one chunk overlaps the host copy of the previous one.
// Returns when the data is in dst.
int copy_to_host(Ctx &c, void *dst, const void *src, size_t bytes) {
    if (bytes == 0) return LSAMD_OK;
    if (bytes <= (size_t(8) << 20)) {
        HIP_OK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c.copy_stream));
        HIP_OK(hipStreamSynchronize(c.copy_stream));
        return LSAMD_OK;
    }
    if (c.pinned_cap < kPinnedChunk) {
        for (int i = 0; i < 2; ++i) {
            if (c.pinned[i]) HIP_OK(hipHostFree(c.pinned[i]));
            c.pinned[i] = nullptr;
            HIP_OK(hipHostMalloc(&c.pinned[i], kPinnedChunk, hipHostMallocDefault));
        }
        c.pinned_cap = kPinnedChunk;
    }
    const size_t nch = (bytes + kPinnedChunk - 1) / kPinnedChunk;
    auto chunk_len = [&](size_t i) { return std::min(kPinnedChunk, bytes - i * kPinnedChunk); };
    const char *s = static_cast<const char *>(src);
    char *d = static_cast<char *>(dst);
    HIP_OK(hipMemcpyAsync(c.pinned[0], s, chunk_len(0), hipMemcpyDeviceToHost, c.copy_stream));
    HIP_OK(hipEventRecord(c.ev[0], c.copy_stream));
    for (size_t i = 1; i <= nch; ++i) {
        if (i < nch) {
            HIP_OK(hipMemcpyAsync(c.pinned[i & 1], s + i * kPinnedChunk, chunk_len(i),
                                  hipMemcpyDeviceToHost, c.copy_stream));
            HIP_OK(hipEventRecord(c.ev[i & 1], c.copy_stream));
        }
        HIP_OK(hipEventSynchronize(c.ev[(i - 1) & 1]));
        const size_t len = chunk_len(i - 1);
        const char *from = static_cast<const char *>(c.pinned[(i - 1) & 1]);
        char *to = d + (i - 1) * kPinnedChunk;
        const size_t parts = 8, step = (len + parts - 1) / parts;
#pragma omp parallel for num_threads(8) schedule(static)
        for (long long p = 0; p < (long long)parts; ++p) {
            const size_t a = (size_t)p * step;
            if (a < len) std::memcpy(to + a, from + a, std::min(step, len - a));
        }
    }
    return LSAMD_OK;
}

double env_double(const char *name, double dflt) {
    const char *v = std::getenv(name);
    if (!v || !*v) return dflt;
    char *end = nullptr;
    const double x = std::strtod(v, &end);
    return (end && end != v) ? x : dflt;
}

}  // namespace
}  // namespace lsamd

using namespace lsamd;

extern "C" {

int lsamd_version(void) { return 10000; }  // 1.0.0

int lsamd_last_error_code(void) { return t_err_code; }
const char *lsamd_last_error(void) { return t_err_msg.c_str(); }
void lsamd_clear_error(void) { clear_error(); }

int lsamd_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

uint32_t lsamd_count_objects(const uint32_t *objectId, uint64_t n) {
    return objectId ? count_runs(objectId, n) : (n ? 1u : 0u);
}

uint32_t lsamd_object_offsets(const uint32_t *objectId, uint64_t n, uint64_t *offsets,
                              uint32_t *ids) {
    if (!objectId || !offsets) return 0;
    return run_offsets(objectId, n, offsets, ids);
}

int lsamd_batch_host(int variant, int dtype, int form, int normalize, int precenter,
                     const uint32_t *objectId, const void *t, void *y, const void *dy, uint64_t n,
                     double fmin, double fmax, uint32_t nf, void *pgram, uint32_t top_k,
                     uint32_t *peak_index, void *peak_power, void *peak_period,
                     uint32_t *num_objects_out) {
    clear_error();
    if (n == 0 || nf == 0) return set_error(LSAMD_ERR_INVALID, "lsamd_batch_host: empty input");
    if (!t || !y || (variant == LSAMD_GENERALIZED && !dy))
        return set_error(LSAMD_ERR_INVALID, "lsamd_batch_host: null input array");
    if (dtype != LSAMD_F32 && dtype != LSAMD_F64)
        return set_error(LSAMD_ERR_INVALID, "lsamd_batch_host: bad dtype");
    if (top_k != 0 && top_k != 1 && top_k != 3)
        return set_error(LSAMD_ERR_INVALID, "lsamd_batch_host: top_k must be 0, 1 or 3");
    if (!(fmax > fmin))
        return set_error(LSAMD_ERR_INVALID, "lsamd_batch_host: max_freq must exceed min_freq");
    const size_t rsz = dtype == LSAMD_F64 ? 8 : 4;

    // segmentation (a NULL id array means one object: reference MODE 2)
    std::vector<uint64_t> off;
    uint32_t K;
    if (objectId) {
        K = count_runs(objectId, n);
        off.resize((size_t)K + 1);
        run_offsets(objectId, n, off.data(), nullptr);
    } else {
        K = 1;
        off = {0, n};
    }
    if (num_objects_out) *num_objects_out = K;

    // reference side effect: the generalized variant pre-centres magY in place
    if (variant == LSAMD_GENERALIZED && precenter) {
        for (uint32_t k = 0; k < K; ++k) {
            if (dtype == LSAMD_F64)
                precenter_host(static_cast<double *>(y) + off[k],
                               static_cast<const double *>(dy) + off[k], off[k + 1] - off[k]);
            else
                precenter_host(static_cast<float *>(y) + off[k],
                               static_cast<const float *>(dy) + off[k], off[k + 1] - off[k]);
        }
    }

    int rc = LSAMD_OK;
    Ctx *cp = ctx_for_current_device(&rc);
    if (!cp) return rc;
    Ctx &c = *cp;
    std::lock_guard<std::mutex> lock(c.mu);
    rc = ctx_init(c);
    if (rc) return rc;

    // memory budget for the periodogram (computeNumBatches, main.cu:423-466):
    // a fraction of the free device memory, LSAMD_MEM_FRACTION (default 0.5),
    // or an absolute LSAMD_PGRAM_BUDGET_BYTES.
    size_t freeb = 0, totalb = 0;
    HIP_OK(hipMemGetInfo(&freeb, &totalb));
    size_t budget = (size_t)(env_double("LSAMD_MEM_FRACTION", 0.5) * (double)freeb);
    const double abs_budget = env_double("LSAMD_PGRAM_BUDGET_BYTES", 0.0);
    if (abs_budget > 0) budget = (size_t)abs_budget;
    const size_t per_obj = pgram ? (size_t)nf * rsz : 0;
    uint32_t objs_per_batch = K;
    if (per_obj > 0) {
        const size_t fit = std::max<size_t>(1, budget / per_obj);
        objs_per_batch = (uint32_t)std::min<size_t>(K, fit);
    }

    // inputs, once
    rc = ensure(c.t, n * rsz);
    if (!rc) rc = ensure(c.y, n * rsz);
    if (!rc && variant == LSAMD_GENERALIZED) rc = ensure(c.dy, n * rsz);
    if (!rc) rc = ensure(c.off, ((size_t)K + 1) * 8);
    if (!rc) rc = ensure(c.order, (size_t)K * 4);
    if (rc) return rc;
    HIP_OK(hipMemcpyAsync(c.t.p, t, n * rsz, hipMemcpyHostToDevice, c.stream));
    HIP_OK(hipMemcpyAsync(c.y.p, y, n * rsz, hipMemcpyHostToDevice, c.stream));
    if (variant == LSAMD_GENERALIZED)
        HIP_OK(hipMemcpyAsync(c.dy.p, dy, n * rsz, hipMemcpyHostToDevice, c.stream));
    HIP_OK(hipMemcpyAsync(c.off.p, off.data(), ((size_t)K + 1) * 8, hipMemcpyHostToDevice,
                          c.stream));

    // Each memory-budget batch is swept in slices of objects: while the
    // periodogram of slice s is copied back on the copy stream, slice s+1 is
    // swept, so the end-to-end time approaches the copy time alone.  Slices
    // hold ~2 GiB of periodogram (at least 1 object; LSAMD_SLICE_BYTES
    // overrides): each slice's copy restarts the two-buffer staging pipeline,
    // so slices much smaller than that lose copy bandwidth.
    const size_t kSliceBytes = (size_t)env_double("LSAMD_SLICE_BYTES", (double)(size_t(2) << 30));
    const uint32_t objs_per_slice =
        per_obj > 0 ? (uint32_t)std::max<size_t>(1, std::min<size_t>(objs_per_batch, kSliceBytes / per_obj))
                    : objs_per_batch;
    std::vector<uint32_t> slice_start;
    for (uint32_t b0 = 0; b0 < K; b0 += objs_per_batch) {
        const uint32_t b1 = std::min(K, b0 + objs_per_batch);
        for (uint32_t s0 = b0; s0 < b1; s0 += objs_per_slice) slice_start.push_back(s0);
    }
    // longest objects first inside every slice (slice-local indices)
    std::vector<uint32_t> order(K);
    for (size_t i = 0; i < slice_start.size(); ++i) {
        const uint32_t s0 = slice_start[i];
        const uint32_t s1 = i + 1 < slice_start.size() ? std::min(slice_start[i + 1], K) : K;
        for (uint32_t k = s0; k < s1; ++k) order[k] = k - s0;
        std::stable_sort(order.begin() + s0, order.begin() + s1, [&](uint32_t a, uint32_t b) {
            return off[s0 + a + 1] - off[s0 + a] > off[s0 + b + 1] - off[s0 + b];
        });
    }
    HIP_OK(hipMemcpyAsync(c.order.p, order.data(), (size_t)K * 4, hipMemcpyHostToDevice,
                          c.stream));

    const uint32_t tk = top_k;
    if (tk) {
        rc = ensure(c.pidx, (size_t)K * tk * 4);
        if (!rc) rc = ensure(c.ppow, (size_t)K * tk * rsz);
        if (!rc) rc = ensure(c.pper, (size_t)K * tk * rsz);
        if (rc) return rc;
    }
    if (pgram) {
        rc = ensure(c.pgram, (size_t)objs_per_batch * per_obj);
        if (rc) return rc;
    }

    // slice i covers objects [s0, s1) and, inside its batch, periodogram rows
    // from (s0 - batch start); a slice never spans two batches
    auto slice_end = [&](size_t i) {
        const uint32_t s0 = slice_start[i];
        const uint32_t batch_end = std::min(K, (s0 / objs_per_batch) * objs_per_batch + objs_per_batch);
        return std::min(batch_end, s0 + objs_per_slice);
    };
    auto batch_row = [&](uint32_t k) { return (size_t)(k % objs_per_batch); };
    auto sweep_slice = [&](size_t i) -> int {
        const uint32_t s0 = slice_start[i], s1 = slice_end(i);
        lsamd_sweep_args a{};
        a.variant = variant;
        a.dtype = dtype;
        a.form = form;
        a.normalize = normalize;
        a.precenter = 0;  // already done on the host, in place
        a.num_objects = s1 - s0;
        a.num_freqs = nf;
        a.top_k = tk;
        a.num_obs = n;
        a.min_freq = fmin;
        a.max_freq = fmax;
        a.offsets = static_cast<const uint64_t *>(c.off.p) + s0;
        a.order = static_cast<const uint32_t *>(c.order.p) + s0;
        a.t = c.t.p;
        a.y = c.y.p;
        a.dy = variant == LSAMD_GENERALIZED ? c.dy.p : nullptr;
        a.pgram = pgram ? static_cast<char *>(c.pgram.p) + batch_row(s0) * per_obj : nullptr;
        // a call of one object may slice its observations; sub-sweeps of a
        // larger call never do, so batching cannot change any result
        a.obs_slices = K == 1 ? 0 : 1;
        if (tk) {
            a.peak_index = static_cast<uint32_t *>(c.pidx.p) + (size_t)s0 * tk;
            a.peak_power = static_cast<char *>(c.ppow.p) + (size_t)s0 * tk * rsz;
            a.peak_period = static_cast<char *>(c.pper.p) + (size_t)s0 * tk * rsz;
        }
        const size_t wsb = lsamd_sweep_workspace_size(&a);
        int r = LSAMD_OK;
        if (c.ws.cap < wsb) {
            // growing the workspace frees it: wait for the sweeps using it
            HIP_OK(hipStreamSynchronize(c.stream));
            r = ensure(c.ws, wsb);
            if (r) return r;
        }
        r = lsamd_sweep(&a, c.ws.p, c.ws.cap, c.stream);
        if (r) return r;
        HIP_OK(hipEventRecord(c.slice[i & 1], c.stream));
        return LSAMD_OK;
    };
    auto copy_slice = [&](size_t i) -> int {
        const uint32_t s0 = slice_start[i], s1 = slice_end(i);
        HIP_OK(hipStreamWaitEvent(c.copy_stream, c.slice[i & 1], 0));
        return copy_to_host(c, static_cast<char *>(pgram) + (size_t)s0 * per_obj,
                            static_cast<char *>(c.pgram.p) + batch_row(s0) * per_obj,
                            (size_t)(s1 - s0) * per_obj);
    };
    const size_t ns = slice_start.size();
    rc = sweep_slice(0);
    if (rc) return rc;
    for (size_t i = 1; i <= ns; ++i) {
        // Within a batch, slice i writes other device rows than slice i-1, so
        // it is swept while slice i-1 is copied back.  The first slice of a
        // new batch reuses rows from 0, which the still-pending last slice of
        // the previous batch may occupy: that slice is copied back first
        // (copy_slice returns once the data is in host memory).
        const bool next = i < ns;
        const bool new_batch = next && slice_start[i] % objs_per_batch == 0;
        if (next && !new_batch) {
            rc = sweep_slice(i);
            if (rc) return rc;
        }
        if (pgram) {
            rc = copy_slice(i - 1);
            if (rc) return rc;
        }
        if (new_batch) {
            rc = sweep_slice(i);
            if (rc) return rc;
        }
    }
    if (tk) {
        if (peak_index)
            HIP_OK(hipMemcpyAsync(peak_index, c.pidx.p, (size_t)K * tk * 4, hipMemcpyDeviceToHost,
                                  c.stream));
        if (peak_power)
            HIP_OK(hipMemcpyAsync(peak_power, c.ppow.p, (size_t)K * tk * rsz,
                                  hipMemcpyDeviceToHost, c.stream));
        if (peak_period)
            HIP_OK(hipMemcpyAsync(peak_period, c.pper.p, (size_t)K * tk * rsz,
                                  hipMemcpyDeviceToHost, c.stream));
    }
    HIP_OK(hipStreamSynchronize(c.stream));
    return LSAMD_OK;
}

int lsamd_cpu_batch(int variant, int dtype, int normalize, const uint32_t *objectId, const void *t,
                    void *y, const void *dy, uint64_t n, double fmin, double fmax, uint32_t nf,
                    void *pgram, int nthreads) {
    clear_error();
    return cpu_batch(variant, dtype, normalize, objectId, t, y, dy, n, fmin, fmax, nf, pgram,
                     nthreads);
}

int lsamd_cpu_single(int variant, int dtype, int normalize, const void *t, void *y, const void *dy,
                     uint64_t n, double fmin, double fmax, uint32_t nf, void *pgram, int nthreads) {
    clear_error();
    return cpu_single(variant, dtype, normalize, t, y, dy, n, fmin, fmax, nf, pgram, nthreads);
}

}  // extern "C"
