/*
 * gx.h -- C ABI of libgx.so, the MI355X (gfx950) cross-section engine.
 *
 * The drop-in boundary is the reference's compute-backend interface,
 * BaseManager (BaseManager.h:6-33), driven by BaseProfile
 * (BaseProfile.cpp:49-74) and the profile drivers (VoigtProfile.cpp:20-26,
 * DopplerProfile.cpp:19-25).  Each gx_* manager entry point below replaces
 * one BaseManager method; gpu_cross_amd/csrc/mi355x_manager.h wraps them in a
 * C++ class with the BaseManager signatures, and gpu_cross_amd/manager.py
 * binds them for Python.
 *
 * Conventions: plain pointers and sizes only; every call returns a gx_status
 * (GX_OK == 0) and leaves a message for gx_last_error() on failure, never
 * exit()s.  Host pointers are borrowed for the duration of the call only.
 * One context drives one GPU from one host thread.
 */
#ifndef GX_H
#define GX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum gx_status {
    GX_OK = 0,
    GX_INVALID_ARG = 1,
    GX_OUT_OF_MEMORY = 2,
    GX_NO_DEVICE = 3,
    GX_DEVICE_ERROR = 4,
    GX_BUSY = 5,
    GX_NOT_FOUND = 6
} gx_status;

/* Input.h:9-15 ProfileType numbering (PSUEDO_VOIGT=3 is not served).  GX_GAUSSIAN is the
 * CUDA manager's fixed-half-width profile (cross_kernal.cu:471-625): gx_consts.half_width is
 * the HWHM, window 10 HWHM, FP64, manager entry points only.  A GX_VOIGT chunk uploaded
 * without gamma and n takes the CUDA manager's sum for that case (cross_kernal.cu:860-922:
 * gamma_L = 0.05, Alg. 916 / 3-point Gauss-Hermite profile, FP64). */
enum { GX_GAUSSIAN = 0, GX_DOPPLER = 1, GX_VOIGT = 2 };
/* Width / window / constant convention: the CPU manager's (OpenMPManager.cpp:146,191,205-206)
 * or the CUDA manager's (cross_kernal.cu:667,938,954-957, GPUManager.cpp:31) */
enum { GX_CONV_OPENMP = 0, GX_CONV_GPU = 1 };
/* FP64 throughout, or FP32 profiles with compensated per-bin sums and FP64
 * intensities / centre offsets (Voigt only; a grid so fine that a line's W4 core
 * spans more than ~4000 bins runs the FP64 path instead) */
enum { GX_PREC_FP64 = 0, GX_PREC_MIXED = 1 };

typedef struct gx_ctx gx_ctx;

/* Arguments of BaseManager::InitializeConstants (BaseManager.h:23) plus the
 * knobs the reference hard-codes. */
typedef struct gx_consts {
    double half_width;      /* Gaussian HWHM in cm^-1 (unused by Doppler/Voigt) */
    double temperature;     /* K */
    double partition;       /* Q(T) */
    double dfreq;           /* bin width of the grid passed to gx_transfer_freq */
    double mean_mass;       /* amu */
    double pressure;        /* atm (gamma arrives already multiplied by P) */
    double ref_temp;        /* K, T_ref of the Lorentz temperature exponent */
    double cutoff;          /* window half-width in cm^-1; <= 0: reference rule
                               (Voigt min(25P,100), Doppler 100) */
    double doppler_widths;  /* > 0: Doppler window k * alpha_i + dfreq per line */
    int convention;         /* GX_CONV_* */
    int precision;          /* GX_PREC_* */
} gx_consts;

/* Kernel times of the last gx_execute / table run, from HIP events on the
 * stream the kernels run on.  A graph replay is timed as a whole (total_ms);
 * its intensity_ms is that of the eager run made before capture and
 * accumulate_ms the remainder. */
typedef struct gx_timing {
    float intensity_ms;     /* line-intensity (and window) kernels */
    float accumulate_ms;    /* bin-parallel profile accumulation kernels */
    float total_ms;         /* whole call */
    long long evals;        /* line-bin evaluations sum_i (ie_i - ib_i) of that call */
    long long lines;        /* lines processed */
    float region1_ms;       /* mixed precision: the W4 region-I accumulation kernel alone
                               (FP64: accumulate_ms) */
    float core_ms;          /* mixed precision: the W4 core pass (regions II-IV); FP64: 0 */
    long long core_evals;   /* mixed precision: evaluations made by the core pass; FP64: 0 */
    float far_ms;           /* mixed precision: the far-field kernels of region I (moments, block
                               plans, far sums); 0 when off */
    long long far_evals;    /* mixed precision: region-I line-bin pairs summed from moments
                               instead of one by one (the region-I kernel made
                               evals - core_evals - far_evals) */
    long long interp_pairs; /* mixed precision: of the region-I kernel's pairs, those formed by
                               interpolation from 16 node evaluations per line and 64-bin
                               block (64 counted per line and block) */
} gx_timing;

/* BaseManager(ProfileType) + GpuManager ctor (GPUManager.cpp:5-17): bind
 * device `device`, take a memory budget (0: device total less 1 GiB). */
int gx_create(int device, int profile, size_t mem_budget, gx_ctx **out);

/* InitializeVectors(int Npoints) + GetNtrans() (GPUManager.cpp:49-76,
 * BaseManager.cpp:25): allocate the grid and report the chunk capacity. */
int gx_initialize_vectors(gx_ctx *ctx, int npoints, size_t *ntrans_out);

/* InitializeConstants(...) (GPUManager.cpp:24-47, OpenMPManager.cpp:57-67). */
int gx_initialize_constants(gx_ctx *ctx, const gx_consts *consts);

/* TransferFreq(h_freq, h_intens, N) (GPUManager.cpp:78-86): grid nodes and
 * the initial intensities.  Blocking. */
int gx_transfer_freq(gx_ctx *ctx, const double *freq, const double *intens, int n);

/* TransferVectors(Nener, E'', nu, A, g, gamma, n) (GPUManager.cpp:87-105):
 * waits for prior work, copies the chunk (gamma/nexp NULL for Doppler and Gaussian;
 * both NULL for a Voigt chunk selects the no-gamma sum), and sorts it by nu on the
 * device when it is not already sorted.  Blocking. */
int gx_transfer_vectors(gx_ctx *ctx, size_t nener, const double *energies, const double *nu, const double *aif,
                        const int *gns, const double *gamma, const double *nexp);

/* ExecuteCrossSection(N, N_ener, start_idx) (GPUManager.cpp:107-121): add
 * the current chunk's first n_ener lines into bins [start_idx, start_idx+points)
 * of the persistent grid.  Asynchronous. */
int gx_execute(gx_ctx *ctx, int points, int n_ener, int start_idx);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): gx_execute on the FP64 path with the per-line
 * intensities, widths and windows, the per-block line slices and the
 * evaluation counters in one fixed device array of the library
 * (GX_NOALLOC_SCRATCH_BYTES on each GPU it runs on, held until the process
 * ends) instead of buffers the context allocates.  Same kernels, so the same
 * grid bit for bit.  The lines and the grid are those the context already
 * holds (gx_transfer_vectors, gx_initialize_vectors).  Refuses more than
 * GX_NOALLOC_MAX_LINES lines, more than GX_NOALLOC_MAX_BINS bins and
 * GX_PREC_MIXED constants (GX_INVALID_ARG).  Calls no hipMalloc or hipFree and
 * issues no memset.  Asynchronous on the context's stream; gx_last_timing does
 * not cover it.  Every context on one GPU shares that GPU's array, which keeps
 * the last call's contents: calls that overlap on one GPU (contexts on other
 * streams, other host threads) race. */
#define GX_NOALLOC_MAX_LINES (1 << 20)
#define GX_NOALLOC_MAX_BINS (1 << 24)
#define GX_NOALLOC_SCRATCH_BYTES ((size_t)34079232) /* 32 B per line, 8 B per 256 bins, 64 counters */
int gx_execute_noalloc(gx_ctx *ctx, int points, int n_ener, int start_idx);

/* ReadyForWork() (GPUManager.cpp:185-188): 1 idle, 0 busy, <0 error. */
int gx_ready(gx_ctx *ctx);

/* TransferResults(h_freq, h_intens, N) (GPUManager.cpp:160-167): waits and
 * overwrites h_intens (and h_freq when not NULL) with the device grid. */
int gx_transfer_results(gx_ctx *ctx, double *freq, double *intens, int n);

/* Cleanup() (GPUManager.cpp:172-183): free device buffers; the context stays valid. */
int gx_cleanup(gx_ctx *ctx);

/* Destroy the context (Cleanup + release of the stream and events). */
int gx_destroy(gx_ctx *ctx);

/* GetNtrans() (BaseManager.cpp:25). */
size_t gx_get_ntrans(gx_ctx *ctx);

/* Block until the context's stream is idle. */
int gx_synchronize(gx_ctx *ctx);

/* Kernel timing of the last gx_execute (synchronises). */
int gx_last_timing(gx_ctx *ctx, gx_timing *out);

/* ------------------------------------------------------------------------
 * Batched (T,P) opacity tables over one resident line list.  No reference
 * counterpart: the reference computes one (T,P) point per process
 * (VoigtProfile.cpp:26).  Each grid of a table equals what the manager path
 * above computes for that point alone.
 * ------------------------------------------------------------------------ */
typedef struct gx_lines gx_lines;
typedef struct gx_table gx_table;

/* Grid rule of BaseProfile.cpp:7-21: dfreq = (end-start)/N, f_i = start + i*dfreq. */
typedef struct gx_grid {
    double start_nu;
    double end_nu;
    int npoints;
} gx_grid;

typedef struct gx_point {
    double temperature;     /* K */
    double pressure;        /* atm (multiplies gamma_ref) */
    double partition;       /* Q(T) */
} gx_point;

typedef struct gx_table_opts {
    int profile;            /* GX_VOIGT, or GX_DOPPLER (FP64 only) */
    int precision;          /* GX_PREC_* */
    int convention;         /* GX_CONV_* */
    double mean_mass;       /* amu */
    double ref_temp;        /* K */
    double cutoff;          /* > 0: fixed; <= 0: reference rule per point */
    double doppler_widths;  /* Doppler: > 0 per-line window k*alpha + dfreq */
    int use_graph;          /* capture the launch sequence once and replay it */
    int autotune;           /* time launch shapes on this table once, keep the fastest */
} gx_table_opts;

/* Upload a line list once (sorted by nu on the device).  gamma_ref is per
 * atm (NULL with nexp NULL: Doppler-only list). */
int gx_lines_create(int device, size_t n, const double *energies, const double *nu, const double *aif,
                    const int *gns, const double *gamma_ref, const double *nexp, gx_lines **out);
int gx_lines_destroy(gx_lines *lines);
size_t gx_lines_count(const gx_lines *lines);

/* Plan a table of npts points on one grid: allocates per-point line data and
 * the npts output grids (admission by free HBM: GX_OUT_OF_MEMORY when the
 * table cannot fit even one point at a time). */
int gx_table_create(gx_lines *lines, const gx_grid *grid, const gx_point *pts, int npts,
                    const gx_table_opts *opts, gx_table **out);
/* Recompute every grid of the table (asynchronous). */
int gx_table_run(gx_table *table);
/* Wait for the last run; timing (may be NULL) gets its kernel times and evaluation count. */
int gx_table_wait(gx_table *table, gx_timing *timing);
/* Copy grid `point` (n bins) to the host. */
int gx_table_read(gx_table *table, int point, double *out, int n);
/* Launch shape in use: bins per lane and line segments per block. */
int gx_table_shape(gx_table *table, int *bins_per_lane, int *segments);
int gx_table_destroy(gx_table *table);

/* ------------------------------------------------------------------------
 * Request server: continuous batching of cross-section requests over one
 * resident line list.  No reference counterpart (the reference is one
 * cross-section per process).  A scheduler thread admits queued requests as
 * in-flight slots and HBM budget free up, co-batches compatible ones (same
 * grid, profile, precision, cutoff) into one launch, runs batches
 * concurrently on separate streams, and records per-request latency.  A
 * request's result does not depend on what it was batched with.
 * ------------------------------------------------------------------------ */
typedef struct gx_server gx_server;

typedef struct gx_request {
    double temperature;     /* K */
    double pressure;        /* atm (multiplies gamma_ref) */
    double partition;       /* Q(T) */
    int profile;            /* GX_VOIGT or GX_DOPPLER */
    int precision;          /* GX_PREC_* (Doppler runs FP64) */
    double start_nu;        /* grid, BaseProfile.cpp:7-21 rule */
    double end_nu;
    int npoints;
    double cutoff;          /* <= 0: reference rule */
    double doppler_widths;  /* Doppler: > 0 per-line window k*alpha + dfreq */
} gx_request;

typedef struct gx_server_opts {
    double mean_mass;       /* amu */
    double ref_temp;        /* K */
    int convention;         /* GX_CONV_* */
    int max_inflight;       /* batches on the device at once (0: 4) */
    int max_batch;          /* requests per batch (0: 16) */
    size_t hbm_budget;      /* bytes for in-flight work (0: free HBM less 4 GiB) */
} gx_server_opts;

enum { GX_REQ_QUEUED = 0, GX_REQ_RUNNING = 1, GX_REQ_DONE = 2, GX_REQ_FAILED = 3, GX_REQ_CANCELLED = 4 };

typedef struct gx_server_summary {
    long long submitted, completed, cancelled, failed, batches, batched_requests;
    long long evals;        /* line-bin evaluations of completed requests */
    double p50_ms, p99_ms, mean_ms, max_ms;   /* submit -> result ready, completed requests */
    double busy_s;          /* wall time from first admission to last completion */
    size_t peak_bytes;      /* largest in-flight device footprint */
    long long shared_requests;   /* requests that ran in a batch of two or more */
} gx_server_summary;

int gx_server_create(gx_lines *lines, const gx_server_opts *opts, gx_server **out);
/* Queue a request; id identifies it in the calls below. */
int gx_server_submit(gx_server *srv, const gx_request *req, uint64_t *id);
/* Current state (GX_REQ_*). */
int gx_server_poll(gx_server *srv, uint64_t id, int *state);
/* Block until the request leaves the queue/device or timeout_s passes (< 0: no limit). */
int gx_server_wait(gx_server *srv, uint64_t id, double timeout_s, int *state);
/* Copy a finished grid (n bins) and release the request; timing may be NULL. */
int gx_server_result(gx_server *srv, uint64_t id, double *out, int n, gx_timing *timing);
/* Cancel a queued or running request (its result is discarded). */
int gx_server_cancel(gx_server *srv, uint64_t id);
int gx_server_stats(gx_server *srv, gx_server_summary *out);
/* Stop the scheduler after the in-flight batches finish and free everything. */
int gx_server_destroy(gx_server *srv);

/* Message of the last failing call on this host thread. */
const char *gx_last_error(void);

/* Library version string. */
const char *gx_version(void);

#ifdef __cplusplus
}
#endif
#endif /* GX_H */
