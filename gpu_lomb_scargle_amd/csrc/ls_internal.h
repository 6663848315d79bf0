// Internal declarations shared by the HIP kernels, the host orchestration
// and the CPU path of the MI355X-native Lomb-Scargle library.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>

#include "lombscargle_amd.h"

namespace lsamd {

// ---- error state (thread-local; ls_host.cpp) -----------------------------
int set_error(int code, const std::string &msg);
void clear_error();

// ---- tile shapes ----------------------------------------------------------
// Frequencies of work item (object, chunk) handled by thread `tid` of a
// workgroup of T threads: f = chunk*T*M + m*T + tid, m = 0..M-1.  Adjacent
// threads own adjacent frequencies, so periodogram stores are coalesced, and
// each thread steps through its M frequencies with a stride of T bins.
struct Tile {
    int threads;  // T
    int fpt;      // M (frequencies per thread)
};
Tile current_tile(int form, int variant, int dtype);
bool tile_supported(int form, int variant, int threads, int fpt);

// ---- workspace layout -----------------------------------------------------
struct Workspace {
    size_t bytes = 0;
    // per object
    size_t obj_shift = 0;   // double[K]  plain generalized pre-centring shift (DTYPE value)
    size_t obj_scale = 0;   // double[K]  power scale (normalisation or 1)
    size_t obj_gen = 0;     // double[8K] split generalized: Wf, Ybar, YY, ybar, W
    // per observation
    size_t rec_y = 0;       // DTYPE[n]   plain: centred y
    size_t rec_w = 0;       // DTYPE[n]   plain generalized: (1/dy)^2 in DTYPE
    size_t rec_t = 0;       // double[n]  split: t
    size_t rec_a = 0;       // float4[n]  split: {y or sqrt(w)(y-ybar), -sin step or sqrt(w), cos, sin step}
    size_t rec_n = 0;       // float[n]   split generalized: -sin step
    size_t rec_r = 0;       // double2[n] split: {cos step, sin step}
    // per (object, chunk) peak partials
    size_t part_pow = 0;    // double[K*chunks*topk]
    size_t part_idx = 0;    // uint32[K*chunks*topk]
    size_t queue = 0;       // uint64   persistent-schedule item counter
    size_t part_sums = 0;   // double[K*chunks*slices*NS*M*T] split forms with slices > 1
    size_t obj_ys = 0;      // float[2K]  matrix-core tile: power-of-two y and weight scales
    size_t colfac = 0;      // matrix-core tile: column-factor fragments, 8 (12 generalized) KiB per 16 observations
    size_t rowfac = 0;      // matrix-core tile: row-step phasors, 4 KiB per 16 observations
    uint32_t chunks = 0;
    uint32_t slices = 1;    // observation slices per (object, chunk)
    Tile tile{};
};
int plan_workspace(const lsamd_sweep_args *a, Workspace *w);

// ---- CPU path (ls_cpu.cpp) -------------------------------------------------
int cpu_batch(int variant, int dtype, int normalize, const uint32_t *objectId, const void *t,
              void *y, const void *dy, uint64_t n, double fmin, double fmax, uint32_t nf,
              void *pgram, int nthreads);
int cpu_single(int variant, int dtype, int normalize, const void *t, void *y, const void *dy,
               uint64_t n, double fmin, double fmax, uint32_t nf, void *pgram, int nthreads);

// ---- segmentation (ls_host.cpp) --------------------------------------------
uint32_t count_runs(const uint32_t *objectId, uint64_t n);
uint32_t run_offsets(const uint32_t *objectId, uint64_t n, uint64_t *offsets, uint32_t *ids);

// ---- reference-semantics host helpers (ls_host.cpp) ------------------------
template <typename R>
void precenter_host(R *y, const R *dy, uint64_t n);

}  // namespace lsamd
