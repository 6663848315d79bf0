/*
 * lombscargle_amd.h -- C ABI of the MI355X-native Lomb-Scargle engine.
 *
 * Two layers:
 *
 *  1. The drop-in entry point LombScarglePy, exported by each of the four
 *     libraries libpyls{noerror,error}{float,double}.so exactly as the
 *     reference exports it (source/main.cu:255-256; library names and flags
 *     from source/Makefile:71-91).  Define LSAMD_DTYPE to float or double
 *     before including this header to get the prototype of the library you
 *     link; it defaults to float, like the reference Python wrapper.
 *
 *  2. The lsamd_* extension API, exported by the same libraries and by
 *     libgpulsamd.so: a stream-ordered sweep over device-resident inputs
 *     (lsamd_sweep), a host-array batch call (lsamd_batch_host), per-object
 *     peak (top-k) outputs without materialising the periodogram, and error
 *     reporting instead of process exit.  The reference has no counterpart
 *     for these beyond the GPU batch path they replace (source/main.cu:742-1280).
 *
 * Everything is plain C: pointers, sizes, enums.  Device pointers are HIP
 * device addresses; `stream` is a hipStream_t passed as void*.
 */
#ifndef LOMBSCARGLE_AMD_H
#define LOMBSCARGLE_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifndef LSAMD_DTYPE
#define LSAMD_DTYPE float
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------
 * Drop-in entry point.  Replaces LombScarglePy, source/main.cu:255-411.
 *
 *   objectId[sizeData]  contiguous runs of equal ids are objects
 *                       (computeObjectRanges, main.cu:468-510)
 *   timeX, magY, magDY  observations; magDY is ignored by the standard
 *                       libraries; the error libraries pre-centre magY IN
 *                       PLACE, as the reference does (main.cu:1631-1662)
 *   minFreq, maxFreq    angular frequencies (rad/day); maxFreq is excluded:
 *                       omega_i = min + i*(max-min)/freqToTest
 *   MODE                1 GPU batch, 2 GPU single object, 4 CPU batch,
 *                       5 CPU single object; other values do nothing
 *   pgram               caller-allocated, K*freqToTest, object-major, fully
 *                       overwritten; the standard libraries normalise it by
 *                       2/(N sigma^2) like the reference Python builds
 *
 * GPU modes compute in the reference's own arithmetic ("plain" form) unless
 * the environment variable LSAMD_FORM selects "split" or "split32".
 * Errors are reported on stderr and through lsamd_last_error*(); the call
 * returns instead of exiting the process (LSAMD_EXIT_ON_ERROR=1 restores
 * the reference's exit).
 * ---------------------------------------------------------------------- */
void LombScarglePy(unsigned int *objectId, LSAMD_DTYPE *timeX, LSAMD_DTYPE *magY,
                   LSAMD_DTYPE *magDY, unsigned int sizeData, double minFreq, double maxFreq,
                   unsigned int freqToTest, int MODE, LSAMD_DTYPE *pgram);

/* ------------------------------------------------------------------------
 * Extension API
 * ---------------------------------------------------------------------- */
enum lsamd_variant { LSAMD_STANDARD = 0, LSAMD_GENERALIZED = 1 };
enum lsamd_dtype { LSAMD_F32 = 0, LSAMD_F64 = 1 };
enum lsamd_form {
    /* the reference's arithmetic: DTYPE phases omega*t, libm-accurate
     * sin/cos, sequential DTYPE sums, tau form; matches the reference CPU
     * path of the same DTYPE */
    LSAMD_FORM_PLAIN = 0,
    /* FP64 phase reduction, FP32 hardware sin/cos seeds with an FP64
     * residual correction and a unit-modulus Newton step, FP64 rotation and
     * three-term recurrence across frequencies, FP64 sums, tau-free FP64
     * power: within ~2e-6 relative of FP64 on the bins that carry power
     * (P >= 1e-3 Pmax), including aliasing frequencies */
    LSAMD_FORM_SPLIT = 1,
    /* as SPLIT but with FP32 arithmetic across frequencies and observations:
     * the default tile (both variants) on the FP16 matrix cores, the sums of
     * 32 x 32 bins as row x column factor products with head/tail splitting
     * and FP32 accumulation, folded every 256 observations (standard: into
     * FP32 sums; generalized: the y sums into FP64, the weight-only sums
     * into FP32); the VALU tiles FP32 rotation / recurrence and FP32 sums
     * over 32 or 64 observations folded into FP64: the fastest form.
     * Bins whose 2x2 condition estimate exceeds 16 (aliases) and objects of
     * at most 16 observations are computed in FP64.  Matrix-core tiles:
     * within 1e-5 of FP64 on every batch measured (7.9e-6 on a 1000-object
     * standard batch, 8.4e-6 generalized); the VALU generalized tiles reach
     * 1.07e-5 on the generalized batch */
    LSAMD_FORM_SPLIT32 = 2
};
enum lsamd_schedule {
    /* one workgroup per (object, frequency chunk) work item, in processing
     * order; the hardware dispatcher balances the load */
    LSAMD_SCHEDULE_GRID = 0,
    /* persistent workgroups (as many as fit on the device) taking the next
     * work item from a counter in the workspace until none are left */
    LSAMD_SCHEDULE_PERSISTENT = 1
};
enum lsamd_status {
    LSAMD_OK = 0,
    LSAMD_ERR_INVALID = 1, /* bad argument */
    LSAMD_ERR_HIP = 2,     /* HIP runtime error */
    LSAMD_ERR_NOMEM = 3,   /* allocation failed */
    LSAMD_ERR_NODEVICE = 4 /* no GPU visible */
};

/* One sweep over device-resident, object-contiguous observations.  Covers
 * the batch kernels (kernel.cu:368-631), the single-object kernels
 * (kernel.cu:76-257; K = 1), the normalisation (main.cu:1319-1352), the
 * generalized pre-centring (main.cu:1631-1662) and the argmax/top-3
 * reductions (main.cu:1136-1219, kernel.cu:5-17), all on the device. */
typedef struct lsamd_sweep_args {
    int32_t variant;      /* enum lsamd_variant */
    int32_t dtype;        /* enum lsamd_dtype: type of t, y, dy, pgram, peak_power, peak_period */
    int32_t form;         /* enum lsamd_form */
    int32_t normalize;    /* standard: scale powers by 2/(N sigma_pop^2) */
    int32_t precenter;    /* generalized, plain form: subtract sum(w y)/sum(w), w=1/sqrt(dy),
                           * per object before the sweep (set 0 if y is already centred) */
    uint32_t num_objects; /* K >= 1 */
    uint32_t num_freqs;   /* Nf >= 1 */
    uint32_t top_k;       /* peaks per object to report: 0, 1 or 3 */
    uint64_t num_obs;     /* total observations n = offsets[K] */
    double min_freq;      /* angular (rad/day) */
    double max_freq;      /* angular, excluded */
    const uint64_t *offsets; /* device, K+1: object k holds rows [offsets[k], offsets[k+1]);
                              * an empty object gets NaN powers (0/0) and peaks at bin 0
                              * with power 0, and leaves the other objects' results as
                              * they are without it */
    const uint32_t *order;   /* device, K: processing order (longest first is best), or NULL */
    const void *t;           /* device, n */
    const void *y;           /* device, n (not modified) */
    const void *dy;          /* device, n; generalized only */
    void *pgram;             /* device, K*Nf, or NULL to skip the periodogram */
    uint32_t *peak_index;    /* device, K*top_k, or NULL */
    void *peak_power;        /* device, K*top_k, or NULL */
    void *peak_period;       /* device, K*top_k (days), or NULL */
    int32_t schedule;        /* enum lsamd_schedule */
    int32_t obs_slices;      /* split forms, 0: a sweep of one object with many observations
                              * spreads them over workgroups (results then differ from the
                              * same object swept in a batch by rounding only); 1: never */
    /* Tile shape of this sweep (threads per workgroup, frequencies per
     * thread), one of the shapes compiled for form/variant (lsamd_autotune
     * reports the fastest); 0, 0: the process default (lsamd_set_tile, else
     * the built-in choice). */
    int32_t tile_threads;
    int32_t tile_fpt;
} lsamd_sweep_args;

/* Bytes of device workspace lsamd_sweep needs for these arguments. */
size_t lsamd_sweep_workspace_size(const lsamd_sweep_args *args);

/* Enqueue the sweep on `stream` (NULL = default stream).  The workspace
 * must be device memory of at least lsamd_sweep_workspace_size() bytes and
 * stay untouched until the sweep completes.  Does not synchronise and does
 * not allocate.  Split forms with few (object, frequency chunk) work items
 * and long objects (e.g. one object) also split the observations over
 * workgroups and add their FP64 sums in a second kernel; the workspace size
 * then includes those per-slice sums. */
int lsamd_sweep(const lsamd_sweep_args *args, void *workspace, size_t workspace_bytes,
                void *stream);

/* lsamd_sweep without allocating and without a workspace argument: the
 * workspace is a static device array of LSAMD_STATIC_SCRATCH_BYTES that the
 * library holds on every GPU it runs on until the process ends.  No
 * counterpart in the reference: for callers that cannot allocate device
 * memory (code under stream capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Arguments as for lsamd_sweep; sweeps whose
 * lsamd_sweep_workspace_size() exceeds the array fail with
 * LSAMD_ERR_INVALID before anything is enqueued.  Every call on one GPU uses
 * that GPU's one array: calls that overlap on one GPU (two streams, two host
 * threads) race and must be serialised by the caller.  Kernel timing
 * (lsamd_set_timing) is not recorded for these calls. */
#define LSAMD_STATIC_SCRATCH_BYTES ((size_t)64 << 20)
int lsamd_sweep_static(const lsamd_sweep_args *args, void *stream);

/* Capture one sweep (all its kernels, on a non-default stream) into a HIP
 * graph for repeated launches with the same arguments and workspace; the
 * capture clears its counters with kernels, not memset nodes. */
typedef struct lsamd_graph lsamd_graph;
int lsamd_graph_create(const lsamd_sweep_args *args, void *workspace, size_t workspace_bytes,
                       void *stream, lsamd_graph **graph);
int lsamd_graph_launch(lsamd_graph *graph, void *stream);
void lsamd_graph_destroy(lsamd_graph *graph);

/* Time every compiled tile shape of args->form/variant/dtype on these
 * arguments (reduced to at most 65536 frequency bins, peaks only) and report
 * the fastest; the caller puts it into args->tile_threads / tile_fpt.  For
 * generalized split32 only the matrix-core tile is a candidate (the VALU
 * tiles reach 1.07e-5 relative on ill-sampled batches, above the form's
 * 1e-5 bound; they remain selectable explicitly).  No process-wide setting
 * changes.  Uses the caller's workspace; synchronises the stream. */
int lsamd_autotune(const lsamd_sweep_args *args, void *workspace, size_t workspace_bytes,
                   void *stream, int *threads, int *freqs_per_thread);
/* Workspace bytes lsamd_autotune needs for these arguments. */
size_t lsamd_autotune_workspace_size(const lsamd_sweep_args *args);

/* ------------------------------------------------------------------------
 * Power at arbitrary frequencies, its derivatives, and peak refinement.
 * The reference has no counterpart: it only evaluates the power on its
 * frequency grid (kernel.cu:484-631, main.cu:1566-1777).  The power is the
 * same function of (omega, t, y, dy) as the reference's -- standard
 * (normalised by 2/(N sigma_pop^2) when `normalize` is set, as its Python
 * builds do) or generalized (AstroPy floating mean, weights dy^-2) -- and is
 * evaluated in FP64 whatever the input dtype.  One wavefront per object.
 * ---------------------------------------------------------------------- */
typedef struct lsamd_point_args {
    int32_t variant;         /* enum lsamd_variant */
    int32_t dtype;           /* enum lsamd_dtype of t, y, dy, grad_y, grad_dy */
    int32_t normalize;       /* standard: scale by 2/(N sigma_pop^2) */
    uint32_t num_objects;    /* K >= 1 */
    uint64_t num_obs;        /* n = offsets[K] */
    const uint64_t *offsets; /* device, K+1 row offsets (each object >= 3 rows) */
    const void *t, *y, *dy;  /* device, n each (dy: generalized only) */
    double *omega;           /* device, K angular frequencies (rad/day); refined in
                              * place by lsamd_refine_peaks */
    double *power;           /* device, K, or NULL: P(omega) */
    double *dpower;          /* device, K, or NULL: dP/domega */
    double *d2power;         /* device, K, or NULL: d2P/domega2 */
    const double *grad_scale;/* device, K, or NULL (= 1): upstream dL/dP per object */
    void *grad_y;            /* device, n, or NULL: grad_scale[k] * dP/dy_j */
    void *grad_dy;           /* device, n, or NULL: grad_scale[k] * dP/d(dy_j) (generalized) */
    /* lsamd_refine_peaks only */
    double max_step;         /* > 0: bound on |step| per Newton iteration (rad/day) */
    double tol;              /* > 0: stop once |step| < tol */
    int32_t max_iter;        /* >= 1 */
    int32_t *iterations;     /* device, K, or NULL: Newton steps taken (negative: not converged) */
} lsamd_point_args;

/* Enqueue: power, dP/domega, d2P/domega2 at omega[k] and, when grad_y /
 * grad_dy are given, the per-observation gradients scaled by grad_scale.
 * dP/dy includes the dependence of the standard normalisation on y. */
int lsamd_power_grad(const lsamd_point_args *args, void *stream);

/* Enqueue: Newton iterations omega <- omega - P'/P'' per object (step bounded
 * by max_step and taken uphill when P'' >= 0), until |step| < tol or max_iter;
 * then power / dpower / d2power at the final omega.  Starting from the grid
 * argmax, this moves each peak below the grid spacing. */
int lsamd_refine_peaks(const lsamd_point_args *args, void *stream);

/* Host-array batch: segmentation by runs of objectId, device staging,
 * lsamd_sweep, copy-back, all synchronous.  y is pre-centred IN PLACE for
 * the generalized variant when precenter is set (reference behaviour).
 * pgram (K*Nf host elements of dtype) may be NULL; peak arrays (K*top_k)
 * may be NULL.  *num_objects_out (may be NULL) receives K. */
int lsamd_batch_host(int variant, int dtype, int form, int normalize, int precenter,
                     const uint32_t *objectId, const void *t, void *y, const void *dy,
                     uint64_t n, double min_freq, double max_freq, uint32_t num_freqs,
                     void *pgram, uint32_t top_k, uint32_t *peak_index, void *peak_power,
                     void *peak_period, uint32_t *num_objects_out);

/* Reference CPU modes 4 and 5 (source/main.cu:1358-1543) with the same
 * arithmetic, OpenMP over objects (batch) or frequencies (single). */
int lsamd_cpu_batch(int variant, int dtype, int normalize, const uint32_t *objectId,
                    const void *t, void *y, const void *dy, uint64_t n, double min_freq,
                    double max_freq, uint32_t num_freqs, void *pgram, int nthreads);
int lsamd_cpu_single(int variant, int dtype, int normalize, const void *t, void *y,
                     const void *dy, uint64_t n, double min_freq, double max_freq,
                     uint32_t num_freqs, void *pgram, int nthreads);

/* Number of runs of equal ids (objects) in objectId[n]. */
uint32_t lsamd_count_objects(const uint32_t *objectId, uint64_t n);
/* Fill offsets[K+1] (row ranges) and ids[K] (may be NULL) of the runs. */
uint32_t lsamd_object_offsets(const uint32_t *objectId, uint64_t n, uint64_t *offsets,
                              uint32_t *ids);

/* Device segmentation (replaces computeObjectRanges, source/main.cu:468-510,
 * for ids already in device memory): offsets[0..K] (device, room for
 * max_objects + 1) of the maximal runs of equal consecutive ids in
 * objectId[n] (device), K in *num_objects (device uint32), and, when order
 * (device, room for max_objects) is not NULL, a processing order for
 * lsamd_sweep: objects by decreasing floor(log2 N_k), ties in index order.
 * If K > max_objects, *num_objects still receives K (saturated at 2^32-1)
 * and nothing else is written.  Stream-ordered; does not allocate; no
 * memset.  The workspace is device memory of
 * lsamd_segment_workspace_size(n, max_objects) bytes. */
size_t lsamd_segment_workspace_size(uint64_t n, uint64_t max_objects);
int lsamd_segment(const uint32_t *objectId, uint64_t n, uint64_t *offsets, uint64_t max_objects,
                  uint32_t *order, uint32_t *num_objects, void *workspace, size_t workspace_bytes,
                  void *stream);

/* Observation table in the reference's input format (importObjXYData,
 * source/main.cu:1781-1843): numbers separated by commas and/or whitespace,
 * rows of objectId, t, y (and dy when with_dy), each value parsed as dtype
 * (LSAMD_F32 / LSAMD_F64) and the id converted from it.  With objectId NULL
 * only *rows is set (size query); otherwise the arrays (capacity rows each)
 * are filled. */
int lsamd_read_table(const char *path, int dtype, int with_dy, uint64_t capacity,
                     uint32_t *objectId, void *t, void *y, void *dy, uint64_t *rows);

/* Best periods to a text file in the reference's formats: top_k = 1 gives
 * "id, period, power" (outputPeriodsToFile, main.cu:961-976), top_k = 3
 * gives "id, p1, p2, p3, pow1, pow2, pow3" (outputPeriodsToFileTopThree,
 * main.cu:979-995); 6 significant digits.  periods/powers hold
 * num_objects*top_k values of dtype. */
int lsamd_write_periods(const char *path, const uint32_t *objectIds, uint32_t num_objects,
                        uint32_t top_k, int dtype, const void *periods, const void *powers);

/* Status of the last failed call on this thread (LSAMD_OK if none). */
int lsamd_last_error_code(void);
const char *lsamd_last_error(void);
void lsamd_clear_error(void);

/* Number of visible HIP devices (0 when none; never fails). */
int lsamd_device_count(void);

/* When enabled for the calling thread, lsamd_sweep and lsamd_graph_launch
 * bracket the sweep kernels with HIP events on their stream;
 * lsamd_last_sweep_kernel_ms() returns the duration of this thread's last
 * bracketed sweep once its stream has been synchronised (negative if
 * unavailable).  The setting and the events are per host thread. */
void lsamd_set_timing(int enabled);
double lsamd_last_sweep_kernel_ms(void);

/* Tile shape of the sweep kernel: threads per workgroup and frequencies per
 * thread, among the shapes compiled for the form and variant (64..512
 * threads, 1..16 frequencies).  SPLIT32 64 x 16 (the built-in choice) is the
 * matrix-core kernel; 256 x 8 (standard) and 256 x 4 (generalized) the
 * fastest VALU ones.  Zero restores the built-in choice.  Returns
 * LSAMD_ERR_INVALID for shapes not compiled.  This is the process-wide
 * default for sweeps whose args leave tile_threads / tile_fpt at 0; a sweep
 * reads it once, when it is planned.  Per-call tiles (the args fields) are
 * the way to run different shapes from several threads. */
int lsamd_set_tile(int form, int variant, int threads, int freqs_per_thread);
/* Tile shape lsamd_sweep uses for args that leave the tile at 0 (for reports). */
void lsamd_get_tile(int form, int variant, int dtype, int *threads, int *freqs_per_thread);

/* Library version (major*10000 + minor*100 + patch). */
int lsamd_version(void);

#ifdef __cplusplus
}
#endif

#endif /* LOMBSCARGLE_AMD_H */
