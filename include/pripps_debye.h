/*
 * pripps_debye.h -- C ABI of the MI355X (gfx950) exact Debye engine.
 *
 * The engine replaces the exact pairwise Debye sum behind the reference's
 * `Pripps::debye_model` (mlund/pripps src/lib.rs:287-303 ->
 * src/debye.rs:78-136).  Everything above that sum -- YAML form factors,
 * structure parsing, solvent dummy construction -- stays on the caller's side
 * of this boundary and arrives here as plain arrays: SoA coordinates, kind
 * ids, and per-(kind, q) form-factor tables in FP64.
 *
 * Conventions
 *   - Every function returns a pripps_status (0 = success).  On failure a
 *     message for the calling thread is available from pripps_last_error().
 *   - Inputs are borrowed for the duration of the call; outputs are
 *     caller-owned arrays.  Handles are opaque.
 *   - A built model is immutable: pripps_debye_model_intensity / _profiles /
 *     _fit may be called concurrently on one handle.  The engine serialises
 *     its own device submissions.
 *   - Profiles use the reference's DebyeTerms order (src/debye.rs:176-184):
 *     0 atoms, 1 excluded, 2 hydration, 3 atom_excluded, 4 atom_hydration,
 *     5 excluded_hydration; layout [6][nq].
 *   - No torch or HIP types cross this boundary.
 */
#ifndef PRIPPS_DEBYE_H
#define PRIPPS_DEBYE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  1..9 mirror the reference's Error variants
 * (src/error.rs:16-74) in declaration order; 10.. are engine-only. */
typedef enum {
    PRIPPS_OK = 0,
    PRIPPS_ERR_FILE = 1,                  /* Error::File */
    PRIPPS_ERR_IO = 2,                    /* Error::Io */
    PRIPPS_ERR_YAML = 3,                  /* Error::Yaml */
    PRIPPS_ERR_UNKNOWN_ATOM = 4,          /* Error::UnknownAtom */
    PRIPPS_ERR_UNDEFINED_FORMFACTOR = 5,  /* Error::UndefinedFormFactor */
    PRIPPS_ERR_Q_OUT_OF_RANGE = 6,        /* Error::QOutOfRange */
    PRIPPS_ERR_VALIDATION = 7,            /* Error::Validation */
    PRIPPS_ERR_USAGE = 8,                 /* Error::Usage */
    PRIPPS_ERR_INTERNAL = 9,              /* Error::Internal */
    PRIPPS_ERR_DEVICE = 10,               /* a HIP runtime call failed */
    PRIPPS_ERR_OUT_OF_MEMORY = 11,        /* device allocation / admission refused */
    PRIPPS_ERR_NOT_READY = 12             /* async request still running */
} pripps_status;

/* Arithmetic of the pair sums. */
typedef enum {
    /* FP32 pair arithmetic, sin(q_k r) by a stable angle-addition recurrence on
     * uniform q grids (direct hardware sin on other grids), FP32 sums over at
     * most `block` pairs carried into FP64 partials, FP64 across tiles. */
    PRIPPS_FP32_COMPENSATED = 0,
    /* The same tiled engine in FP64 throughout. */
    PRIPPS_FP64 = 1,
    /* Plain FP64: one libm-style sin per pair and q, no recurrence, no tiling
     * beyond LDS staging -- the straightforward GPU transcription of the
     * reference loop, kept as a baseline. */
    PRIPPS_FP64_PLAIN = 2
} pripps_precision;

/* One structure.  Coordinates in Angstrom, q in 1/Angstrom. */
typedef struct {
    int64_t n_atoms;
    const double *x, *y, *z;           /* [n_atoms] SoA positions */
    const int32_t *kind;               /* [n_atoms] ids in [0, n_kinds) */
    int32_t n_kinds;
    int32_t nq;
    const double *q;                   /* [nq] q values, >= 0 */
    const double *atom_ff;             /* [n_kinds][nq] F_k(q_m) of the atoms */
    const double *excluded_ff;         /* [n_kinds][nq] excluded-volume dummy form
                                          factor at unit volume scale, or NULL
                                          (excluded volume disabled) */
    const double *water_ff;            /* [nq] shared hydration form factor, or NULL */
    const double *hydration_weight;    /* [n_atoms] per-atom hydration weight, or
                                          NULL (hydration disabled) */
} pripps_debye_input;

typedef struct pripps_engine pripps_engine;
typedef struct pripps_debye_model pripps_debye_model;

/* Kernel tuning; zero fields mean "choose automatically". */
typedef struct {
    int32_t q_per_lane;   /* q points per lane strip: 32, 64 or 128 (FP64: 32, 64);
                             256: the Chebyshev kernel (FP32 or FP64, strips of up
                             to 256 q-points, any grid; the default from 3e7
                             pair-q terms, N^2 nq / 2) */
    int32_t block;        /* pairs summed in FP32 before the FP64 carry */
    int32_t span;         /* atoms per work item along j; Chebyshev plans: j atoms per
                             work unit (0: one unit per wave slot of the GPU) */
    int32_t variant;      /* FP32 inner-loop variant for the autotuner (0 = default) */
} pripps_tuning;

/* Result of a fit (reference FitResult, src/fit.rs:238-247). */
typedef struct {
    double volume_scale;
    double contrast_density;
    double chi2;
    double scale;
} pripps_fit_result;

typedef enum { PRIPPS_FIT_NELDER_MEAD = 0, PRIPPS_FIT_GRID = 1 } pripps_fit_method;

/* Timing / work counters of the last build on a model. */
typedef struct {
    double kernel_ms;            /* device time of the pair kernels */
    double total_ms;             /* host wall time of the build call */
    double pair_q_terms;         /* N(N-1)/2 * nq */
    int32_t q_per_lane, block, span;  /* span 0: Chebyshev units sized automatically */
    int64_t work_items;          /* queue entries: items, or units of a Chebyshev plan */
    int32_t frames;              /* frames averaged (pripps_debye_frames_average), else 0 */
    int32_t frames_per_launch;   /* frames per persistent launch */
    int32_t graph_launches;      /* chunks submitted by replaying the captured HIP graph */
    int32_t cheb_order;          /* expansion order of a Chebyshev plan (q_per_lane 256), else 0 */
} pripps_build_stats;

/* Thread-local message of the last failed call on this thread ("" if none). */
const char *pripps_last_error(void);

/* Library version string. */
const char *pripps_version(void);

/* Engine on HIP device `device`.  Owns a stream and device scratch. */
int pripps_engine_create(int device, pripps_engine **out);
int pripps_engine_destroy(pripps_engine *engine);
/* Free / total device memory as seen by the engine (bytes). */
int pripps_engine_memory(pripps_engine *engine, uint64_t *free_bytes, uint64_t *total_bytes);

/*
 * Build a cached Debye model: the six partial profiles of `input`, kept on
 * the device.  Replaces debye::build_model (src/debye.rs:78-136) called from
 * Pripps::debye_model (src/lib.rs:287-303).  `tuning` may be NULL.
 */
int pripps_debye_model_build(pripps_engine *engine, const pripps_debye_input *input,
                             int32_t precision, const pripps_tuning *tuning,
                             pripps_debye_model **out);

/*
 * Build many models in one persistent launch per kernel shape (structures of
 * uneven size and q grid share one device work queue; FP32 Chebyshev plans
 * run one launch per expansion order).  `out` receives n handles.  The
 * reference has no batched entry point; its multi-frame driver is the serial
 * loop of trajectory_average (src/trajectory.rs:167-251).
 */
int pripps_debye_model_build_batch(pripps_engine *engine, int32_t n,
                                   const pripps_debye_input *inputs, int32_t precision,
                                   const pripps_tuning *tuning, pripps_debye_model **out);

/* Number of q points (DebyeModel::q_values length, src/debye.rs:171-173). */
int pripps_debye_model_nq(const pripps_debye_model *model, int32_t *nq);
/* The q grid the model was built over (src/debye.rs:171-173). */
int pripps_debye_model_q_values(const pripps_debye_model *model, double *out);

/*
 * I(q) at (volume_scale, contrast_density): total plus the unscaled self
 * terms, DebyeModel::intensity (src/debye.rs:146-168) with
 * DebyeTerms::combine (src/debye.rs:186-195).  Any output may be NULL.
 */
int pripps_debye_model_intensity(const pripps_debye_model *model, double volume_scale,
                                 double contrast_density, double *total, double *atoms,
                                 double *excluded, double *hydration);

/* The six cached profiles, [6][nq] (the reference caches them in DebyeTerms
 * but does not expose the cross terms). */
int pripps_debye_model_profiles(const pripps_debye_model *model, double *out);

/*
 * Fit (volume_scale, contrast_density) to experimental data on the device
 * (FP64): reference fit_parameters_with_bounds (src/fit.rs:365-411) and
 * fit_parameters_grid_with_bounds (src/fit.rs:446-516) over the closure
 * `model.intensity` (pripps-py/src/lib.rs:625-650, 781-803).  `sigma` may be
 * NULL (sigma = 1).  Bounds: pointers to (lo, hi) or NULL for the defaults
 * (0.5, 2.0) and (0.0, 0.1).
 */
int pripps_debye_model_fit(const pripps_debye_model *model, const double *q,
                           const double *intensity, const double *sigma, int64_t n,
                           int32_t method, int32_t c1_points, int32_t c2_points,
                           const double *c1_bounds, const double *c2_bounds,
                           pripps_fit_result *out);

/*
 * The fitted theory curve at experimental q: the natural cubic spline of the
 * combined total I(q; volume_scale, contrast_density) over the model's grid,
 * evaluated at q[0..n) (constant beyond the ends) -- the i_fit column of
 * write_fit_csv (src/fit.rs:548-560) before its scale factor.  FP64.
 */
int pripps_debye_model_fit_curve(const pripps_debye_model *model, double volume_scale,
                                 double contrast_density, const double *q, int64_t n,
                                 double *out);

/*
 * A model from previously exported profiles ([6][nq], pripps_debye_model_profiles)
 * and their q grid: checkpoint / resume of a cached model without rebuilding.
 * The reference's DebyeModel is not serialisable (src/debye.rs:139-142); its
 * fit functions take any (c1, c2) -> Intensity closure (src/fit.rs:365-516),
 * and a model restored from chosen profiles plays that role here.
 */
int pripps_debye_model_from_profiles(pripps_engine *engine, const double *q, int32_t nq,
                                     const double *profiles, pripps_debye_model **out);

int pripps_debye_model_stats(const pripps_debye_model *model, pripps_build_stats *out);

/*
 * Re-run the pair and finalize kernels of the engine's most recent build
 * `repeats` times on its device-resident inputs (no host work, no upload).
 * With `pair_ms_mean` non-NULL the kernels are launched one by one and
 * `*pair_ms_mean` is the mean device time of the pair kernel per run (HIP
 * events on the engine stream).  With `pair_ms_mean` NULL each run is one
 * launch of a HIP graph that holds the build's kernel launches (captured at
 * the first such call after a build, kept until the next build).
 * `*total_ms` (if non-NULL): device time of all runs.  `model` must belong to
 * that build (any model of a batch); otherwise PRIPPS_ERR_USAGE.  Results
 * are bitwise those of the build.  No reference counterpart: it exists to
 * measure the kernels and replay them with inputs in HBM.
 */
int pripps_debye_model_recompute(pripps_debye_model *model, int32_t repeats,
                                 double *pair_ms_mean, double *total_ms);
int pripps_debye_model_free(pripps_debye_model *model);

/*
 * Build without device allocation.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Computes
 * the six profiles of one structure like pripps_debye_model_build with the
 * same plan (same tuning, or the default plan), but every device buffer it
 * needs beyond the caller's arrays is one fixed device array of
 * PRIPPS_NOALLOC_SCRATCH_BYTES per GPU, held by the library until the
 * process ends; the call makes no hipMalloc, hipFree or other allocating
 * device call.  `profiles_out` (host, [6][nq], DebyeTerms order) receives
 * the profiles; no model is created.  Without an explicit tuning span, a
 * plan whose work lists do not fit is built with longer j-spans (fewer items
 * or units) instead, which changes the FP32 summation order but not the
 * precision bound; a structure that does not fit even so is refused with
 * PRIPPS_ERR_OUT_OF_MEMORY (Chebyshev plans on a uniform grid of 500 q:
 * FP32 120,000 globule atoms fit, 125,000 do not; FP64 110,000 fit, 120,000
 * do not).  Calls
 * on one GPU share the array and are serialised by the library.  Precision:
 * FP32_COMPENSATED or FP64.
 */
#define PRIPPS_NOALLOC_SCRATCH_BYTES (64u << 20)
int pripps_debye_build_noalloc(pripps_engine *engine, const pripps_debye_input *input,
                               int32_t precision, const pripps_tuning *tuning,
                               double *profiles_out);

/* ---------------------------------------------------------------------- */
/* Autotuning.  No reference counterpart (the reference has no kernels).   */
/* ---------------------------------------------------------------------- */
/* One entry of an engine's tuning table: the plan used for structures of
 * about 2^log2_atoms atoms and nq <= 2^log2_nq at one precision and q-grid
 * kind.  span_shift scales the default j-span by 2^span_shift. */
typedef struct {
    int32_t precision;
    int32_t uniform;       /* 1: uniform q grid (angle-addition recurrence) */
    int32_t log2_atoms;    /* round(log2(n_atoms)) */
    int32_t log2_nq;       /* ceil(log2(nq)) */
    int32_t q_per_lane;
    int32_t block;
    int32_t span_shift;
    int32_t variant;
    double pair_ms;        /* measured pair-kernel time of the tuned input */
} pripps_tuning_entry;

/*
 * Time candidate plans (q points per lane, j-span, pair block, inner-loop
 * variant) on `input` by coordinate descent, `repeats` timed passes each,
 * and record the fastest in the engine's tuning table.  Later builds whose
 * `tuning` is NULL (or all zero) take the table entry of their size class;
 * without one they use the built-in defaults.  `best` (may be NULL) receives
 * the winning plan in absolute terms for this input.  A structure's plan
 * depends only on the structure and the table, so its profiles stay the same
 * alone and in a batch.
 */
int pripps_engine_autotune(pripps_engine *engine, const pripps_debye_input *input,
                           int32_t precision, int32_t repeats, pripps_tuning *best,
                           double *best_pair_ms);
/* Copy up to `capacity` table entries to `entries`; *count = entries held. */
int pripps_engine_tuning_table(pripps_engine *engine, pripps_tuning_entry *entries,
                               int32_t capacity, int32_t *count);
/* Insert or replace the entry of one size class (e.g. a saved table). */
int pripps_engine_set_tuning_entry(pripps_engine *engine, const pripps_tuning_entry *entry);

/* ---------------------------------------------------------------------- */
/* Trajectory frames.                                                      */
/* ---------------------------------------------------------------------- */
/*
 * Weighted average of the six profiles over the frames of a trajectory:
 * trajectory_average (src/trajectory.rs:167-251) with the stride / weight
 * semantics of XtcIterator (src/trajectory.rs:29-135).  `topology` gives the
 * kinds, q grid, form-factor tables and hydration weights; its x/y/z are
 * not read.  `coords` holds [n_frames][n_atoms][3] positions in Angstrom.
 * Frames 0, stride, 2 stride, ... are used; `frame_weights` ([n_frames],
 * indexed by the original frame number, NULL = all ones) must have a finite
 * non-zero sum, and so must the weights of the frames used
 * (PRIPPS_ERR_VALIDATION otherwise).  The result's profiles are
 * sum_f w_f P_f / sum_f w_f, so its intensity(vs, cd) is the weighted average
 * of the per-frame intensities.  Frames run in chunks of up to
 * max_frames_per_launch (0 = as many as a 4 GiB workspace holds) through the
 * batched persistent launch; the first full chunk's submission (upload, pair
 * kernels, reduction, weighted accumulation) is captured as a HIP graph and
 * every further full chunk replays it.  Precision: FP32_COMPENSATED or FP64.
 */
int pripps_debye_frames_average(pripps_engine *engine, const pripps_debye_input *topology,
                                int32_t n_frames, const double *coords,
                                const double *frame_weights, int32_t stride, int32_t precision,
                                const pripps_tuning *tuning, int32_t max_frames_per_launch,
                                pripps_debye_model **out);

/* ---------------------------------------------------------------------- */
/* Serving: continuous batching with admission by free device memory.      */
/* ---------------------------------------------------------------------- */
typedef struct pripps_server pripps_server;

typedef struct {
    int32_t precision;          /* pripps_precision of every request */
    int32_t max_in_flight;      /* requests on the device at once (0 = auto) */
    uint64_t memory_budget;     /* bytes of device memory for requests (0 = 90% of free) */
} pripps_server_config;

int pripps_server_create(pripps_engine *engine, const pripps_server_config *config,
                         pripps_server **out);
/* Queue a build; returns at once with a request id.  The input arrays are
 * copied before the call returns. */
int pripps_server_submit(pripps_server *server, const pripps_debye_input *input,
                         int64_t *request_id);
/* Wait up to timeout_ms (negative = forever) for a request.  On success the
 * caller owns *model.  PRIPPS_ERR_NOT_READY when the wait timed out.  A
 * request's status is its own: when a batch fails, its requests are built
 * again one by one, so one request with invalid input (a non-finite
 * coordinate, say) fails alone.  One call at a time may wait for a request; a
 * second concurrent wait for the same id returns PRIPPS_ERR_USAGE.  After a
 * successful or failed (not timed-out) wait the id is unknown. */
int pripps_server_wait(pripps_server *server, int64_t request_id, int32_t timeout_ms,
                       pripps_debye_model **model, double *latency_ms);
/* Stops the lanes after the queued requests are built, then frees the
 * requests and models nobody waited for.  Call it after every wait has
 * returned. */
int pripps_server_destroy(pripps_server *server);

#ifdef __cplusplus
}
#endif

#endif /* PRIPPS_DEBYE_H */
