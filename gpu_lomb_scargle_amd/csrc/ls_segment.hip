// Device segmentation of an observation stream into objects
// (lsamd_segment): the runs of equal ids become row offsets, and a processing
// order groups objects by decreasing length class.
//
// Replaces the host-side computeObjectRanges (source/main.cu:468-510) for
// callers whose ids are already in device memory.  The run rule is the
// reference's: an object is a maximal run of equal consecutive ids, so an id
// that reappears later starts a new object.
//
// Offsets: block-level head counts, one workgroup scanning the block counts,
// then every block writes the positions of its heads at its scanned base.
// Order: a stable counting sort of objects by floor(log2 N_k), largest class
// first (exact length order is not needed to balance the sweep's work
// queue, only longest-first at the granularity of a factor of two).

#include <hip/hip_runtime.h>

#include <string>

#include "ls_internal.h"

namespace lsamd {
namespace {

constexpr int kSegThreads = 256;
constexpr int kSegItems = 8;                       // ids per thread
constexpr int kSegTile = kSegThreads * kSegItems;  // ids per block
constexpr int kClasses = 33;                       // floor(log2 N) for N in [1, 2^32]

__device__ __forceinline__ bool is_head(const uint32_t *id, uint64_t i) {
    return i == 0 || id[i] != id[i - 1];
}

// Exclusive prefix of v over the block (256 threads = 4 waves); returns the
// block total in *total.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t *s_wave, uint32_t *total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) s_wave[wave] = x;
    __syncthreads();
    uint32_t base = 0, sum = 0;
#pragma unroll
    for (int w = 0; w < kSegThreads / 64; ++w) {
        if (w < wave) base += s_wave[w];
        sum += s_wave[w];
    }
    __syncthreads();
    *total = sum;
    return base + x - v;
}

__global__ __launch_bounds__(kSegThreads) void k_seg_count(const uint32_t *id, uint64_t n,
                                                           uint32_t *block_counts) {
    __shared__ uint32_t s_wave[kSegThreads / 64];
    const uint64_t base = (uint64_t)blockIdx.x * kSegTile + (uint64_t)threadIdx.x * kSegItems;
    uint32_t c = 0;
#pragma unroll
    for (int k = 0; k < kSegItems; ++k) {
        const uint64_t i = base + k;
        if (i < n && is_head(id, i)) ++c;
    }
    uint32_t total;
    block_exclusive_scan(c, s_wave, &total);
    if (threadIdx.x == 0) block_counts[blockIdx.x] = total;
}

// One workgroup: exclusive scan of nb counts in place (64-bit bases), K.
__global__ __launch_bounds__(kSegThreads) void k_seg_scan(const uint32_t *block_counts, uint64_t nb,
                                                          uint64_t *block_base, uint64_t n,
                                                          uint64_t *offsets, uint64_t max_objects,
                                                          uint32_t *num_objects, uint32_t *overflow) {
    __shared__ uint32_t s_wave[kSegThreads / 64];
    uint64_t carry = 0;
    for (uint64_t b0 = 0; b0 < nb; b0 += kSegThreads) {
        const uint64_t b = b0 + threadIdx.x;
        const uint32_t v = b < nb ? block_counts[b] : 0u;
        uint32_t total;
        const uint32_t ex = block_exclusive_scan(v, s_wave, &total);
        if (b < nb) block_base[b] = carry + ex;
        carry += total;
    }
    if (threadIdx.x == 0) {
        const bool fits = carry <= max_objects;
        *overflow = fits ? 0u : 1u;
        *num_objects = carry > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)carry;
        if (fits) offsets[carry] = n;
    }
}

__global__ __launch_bounds__(kSegThreads) void k_seg_write(const uint32_t *id, uint64_t n,
                                                           const uint64_t *block_base, uint64_t *offsets,
                                                           const uint32_t *overflow) {
    __shared__ uint32_t s_wave[kSegThreads / 64];
    if (*overflow) return;  // uniform across the grid
    const uint64_t base = (uint64_t)blockIdx.x * kSegTile + (uint64_t)threadIdx.x * kSegItems;
    uint32_t c = 0;
#pragma unroll
    for (int k = 0; k < kSegItems; ++k) {
        const uint64_t i = base + k;
        if (i < n && is_head(id, i)) ++c;
    }
    uint32_t total;
    uint64_t pos = block_base[blockIdx.x] + block_exclusive_scan(c, s_wave, &total);
#pragma unroll
    for (int k = 0; k < kSegItems; ++k) {
        const uint64_t i = base + k;
        if (i < n && is_head(id, i)) offsets[pos++] = i;
    }
}

__device__ __forceinline__ int length_rank(const uint64_t *offsets, uint32_t k) {
    const uint64_t len = offsets[k + 1] - offsets[k];  // >= 1
    int cls = 63 - __clzll((long long)len);              // floor(log2 len)
    cls = cls < kClasses - 1 ? cls : kClasses - 1;
    return kClasses - 1 - cls;                           // largest class first
}

// Per block of objects: count per length rank -> counts[rank * nbo + block].
__global__ __launch_bounds__(kSegThreads) void k_ord_count(const uint64_t *offsets, const uint32_t *num_objects,
                                                           const uint32_t *overflow, uint32_t nbo,
                                                           uint32_t *counts) {
    __shared__ uint32_t s_c[kClasses];
    if (threadIdx.x < kClasses) s_c[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t K = *overflow ? 0u : *num_objects;  // K <= max_objects when set
    const uint32_t k = blockIdx.x * kSegThreads + threadIdx.x;
    if (k < K) atomicAdd(&s_c[length_rank(offsets, k)], 1u);
    __syncthreads();
    if (threadIdx.x < kClasses) counts[threadIdx.x * nbo + blockIdx.x] = s_c[threadIdx.x];
}

// One workgroup: exclusive scan of the rank-major count matrix in place.
__global__ __launch_bounds__(kSegThreads) void k_ord_scan(uint32_t *counts, uint64_t total_cells) {
    __shared__ uint32_t s_wave[kSegThreads / 64];
    uint32_t carry = 0;
    for (uint64_t c0 = 0; c0 < total_cells; c0 += kSegThreads) {
        const uint64_t c = c0 + threadIdx.x;
        const uint32_t v = c < total_cells ? counts[c] : 0u;
        uint32_t total;
        const uint32_t ex = block_exclusive_scan(v, s_wave, &total);
        if (c < total_cells) counts[c] = carry + ex;
        carry += total;
    }
}

// Stable placement: rank within the block among objects of the same class.
__global__ __launch_bounds__(kSegThreads) void k_ord_write(const uint64_t *offsets, const uint32_t *num_objects,
                                                           const uint32_t *overflow, uint32_t nbo,
                                                           const uint32_t *bases, uint32_t *order) {
    __shared__ int s_rank[kSegThreads];
    const uint32_t K = *overflow ? 0u : *num_objects;
    const uint32_t k = blockIdx.x * kSegThreads + threadIdx.x;
    const int r = k < K ? length_rank(offsets, k) : -1;
    s_rank[threadIdx.x] = r;
    __syncthreads();
    if (k >= K) return;
    uint32_t before = 0;
    for (uint32_t j = 0; j < threadIdx.x; ++j) before += (s_rank[j] == r) ? 1u : 0u;
    order[bases[(uint32_t)r * nbo + blockIdx.x] + before] = k;
}

}  // namespace

#define LS_HIP(call)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess)                                                                \
            return set_error(LSAMD_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

namespace {
struct SegLayout {
    size_t block_counts, block_base, overflow, ord_counts, bytes;
};
size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }
SegLayout seg_layout(uint64_t n, uint64_t max_objects) {
    SegLayout l{};
    const uint64_t nb = (n + kSegTile - 1) / kSegTile;
    const uint64_t nbo = (max_objects + kSegThreads - 1) / kSegThreads;
    size_t off = 0;
    l.block_counts = off;
    off = align256(off + 4 * nb);
    l.block_base = off;
    off = align256(off + 8 * nb);
    l.overflow = off;
    off = align256(off + 4);
    l.ord_counts = off;
    off = align256(off + 4 * (size_t)kClasses * nbo);
    l.bytes = off + 256;
    return l;
}
}  // namespace

}  // namespace lsamd

using namespace lsamd;

extern "C" {

size_t lsamd_segment_workspace_size(uint64_t n, uint64_t max_objects) {
    return seg_layout(n, max_objects).bytes;
}

int lsamd_segment(const uint32_t *objectId, uint64_t n, uint64_t *offsets, uint64_t max_objects,
                  uint32_t *order, uint32_t *num_objects, void *workspace, size_t workspace_bytes,
                  void *stream_v) {
    if (!objectId || !offsets || !num_objects || n == 0)
        return set_error(LSAMD_ERR_INVALID, "lsamd_segment: null pointer or empty input");
    if (max_objects == 0 || max_objects > 0xFFFFFFFFull)
        return set_error(LSAMD_ERR_INVALID, "lsamd_segment: max_objects must be in 1..2^32-1");
    const SegLayout l = seg_layout(n, max_objects);
    if (!workspace || workspace_bytes < l.bytes)
        return set_error(LSAMD_ERR_INVALID, "lsamd_segment: workspace too small (need " +
                                                std::to_string(l.bytes) + " bytes)");
    char *ws = reinterpret_cast<char *>((reinterpret_cast<uintptr_t>(workspace) + 255) & ~(uintptr_t)255);
    uint32_t *block_counts = reinterpret_cast<uint32_t *>(ws + l.block_counts);
    uint64_t *block_base = reinterpret_cast<uint64_t *>(ws + l.block_base);
    uint32_t *overflow = reinterpret_cast<uint32_t *>(ws + l.overflow);
    uint32_t *ord_counts = reinterpret_cast<uint32_t *>(ws + l.ord_counts);
    hipStream_t st = static_cast<hipStream_t>(stream_v);
    const uint64_t nb = (n + kSegTile - 1) / kSegTile;
    if (nb > 0x7FFFFFFFull) return set_error(LSAMD_ERR_INVALID, "lsamd_segment: input too large");
    hipLaunchKernelGGL(k_seg_count, dim3((uint32_t)nb), dim3(kSegThreads), 0, st, objectId, n, block_counts);
    LS_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_seg_scan, dim3(1), dim3(kSegThreads), 0, st, block_counts, nb, block_base, n, offsets,
                       max_objects, num_objects, overflow);
    LS_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_seg_write, dim3((uint32_t)nb), dim3(kSegThreads), 0, st, objectId, n, block_base,
                       offsets, overflow);
    LS_HIP(hipGetLastError());
    if (order) {
        // on overflow (K > max_objects) the order kernels see K = 0
        const uint32_t nbo = (uint32_t)((max_objects + kSegThreads - 1) / kSegThreads);
        hipLaunchKernelGGL(k_ord_count, dim3(nbo), dim3(kSegThreads), 0, st, offsets, num_objects, overflow,
                           nbo, ord_counts);
        LS_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_ord_scan, dim3(1), dim3(kSegThreads), 0, st, ord_counts,
                           (uint64_t)kClasses * nbo);
        LS_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_ord_write, dim3(nbo), dim3(kSegThreads), 0, st, offsets, num_objects, overflow,
                           nbo, ord_counts, order);
        LS_HIP(hipGetLastError());
    }
    return LSAMD_OK;
}

}  // extern "C"
