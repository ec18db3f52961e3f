/*
 * mp_gp_backend.h — C ABI of the MI355X (gfx950) global-placement numeric backend.
 *
 * Plain C types only: pointers, sizes, doubles.  Every entry point returns an
 * mp_gp_status; the message of the last failure on the calling thread is
 * returned by mp_gp_last_error().
 *
 * Reference interfaces replaced (paths relative to the MyPlacement tree):
 *   class CudaPlacementBackend            src/placement/CudaPlacementBackend.hpp:25-39
 *   tryCreateCudaPlacementBackend         src/placement/CudaPlacementBackend.hpp:44-48
 *   cudaPlacementBackendCompiled          src/placement/CudaPlacementBackend.hpp:44
 *   CudaPlacementEvaluation               src/placement/CudaPlacementBackend.hpp:15-23
 *   DensityMetrics                        include/myplacement/metrics/Metrics.hpp:38-47
 *   HostLayout / buildHostLayout          src/placement/cuda/CudaPlacementBackend.cu:764-839
 *   GlobalPlacementOptions (adaptive)     include/myplacement/placement/GlobalPlacer.hpp:26-80
 *   GlobalPlacementIteration / Result     include/myplacement/placement/GlobalPlacer.hpp:82-129
 *   runAdaptiveGlobalPlacement            src/placement/AdaptiveGlobalPlacer.cpp:652-953
 *   randomPlacement / quadraticPlacement  src/placement/InitialPlacer.cpp:102-120, 266-402
 */
#ifndef MP_GP_BACKEND_H
#define MP_GP_BACKEND_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  They map one-to-one onto the reference's error behaviour:
 *   MP_GP_INVALID_ARGUMENT  <-> std::invalid_argument  (.cu:923-928, AdaptiveGlobalPlacer.cpp:186-187)
 *   MP_GP_RUNTIME_ERROR     <-> std::runtime_error     (.cu:30-43, .cu:945-948)
 *   MP_GP_UNAVAILABLE       <-> factory returns nullptr with `reason` set (.cu:1156-1216)           */
typedef enum mp_gp_status {
    MP_GP_OK = 0,
    MP_GP_INVALID_ARGUMENT = 1,
    MP_GP_RUNTIME_ERROR = 2,
    MP_GP_UNAVAILABLE = 3
} mp_gp_status;

/* Particle layers of the density scatter (.cu:811-822): standard cell, movable macro, filler. */
enum { MP_LAYER_STANDARD = 0, MP_LAYER_MACRO = 1, MP_LAYER_FILLER = 2 };

/* Arithmetic used for the DCT bin-grid products: FP32 MFMA, or split-BF16 (each operand as
 * hi + lo bf16 pair, three bf16 MFMA products accumulated in FP32, relative error ~1e-5 per
 * product); MP_DCT_BF16X3 needs both bin counts to be multiples of 64. */
enum { MP_DCT_FP32 = 0, MP_DCT_BF16X3 = 1 };

/* ---- netlist description (the PlacementDatabase fields the GP path reads) ---------------
 * Mirrors include/myplacement/model/PlacementDatabase.hpp:25-71.  Pins are listed in net
 * order: the pins of net n are [net_pin_start[n], net_pin_start[n+1]), which is the pin-id
 * order BookshelfParser produces (src/io/BookshelfParser.cpp:244-271).                     */
typedef struct mp_design {
    int64_t module_count;
    const double* width;          /* [M] */
    const double* height;         /* [M] */
    const double* center;         /* [2M] x,y: Module::center */
    const int32_t* orientation;   /* [M] N,S,E,W,FN,FS,FE,FW,Unknown = 0..8 */
    const uint8_t* is_terminal;   /* [M] */
    const uint8_t* is_fixed;      /* [M] as parsed (terminals are forced fixed on refresh) */
    int64_t net_count;
    const double* net_weight;     /* [N] */
    const int64_t* net_pin_start; /* [N+1] */
    int64_t pin_count;
    const int32_t* pin_module;    /* [I] */
    const double* pin_offset;     /* [2I] raw (un-oriented) offsets from the module centre */
    int64_t row_count;
    const double* rows;           /* [5R] bottom, height, x_start, site_spacing, site_count */
} mp_design;

/* ---- flattened backend input: HostLayout + static density layers ------------------------
 * Mirrors HostLayout (.cu:764-775) and the static upload (.cu:897-917).                    */
typedef struct mp_gp_static {
    int64_t particle_count;              /* movable modules (movableModules() order), then fillers */
    const double* particle_width;        /* [P] */
    const double* particle_height;       /* [P] */
    const double* particle_charge_area;  /* [P] densityChargeArea (GlobalPlacementSupport.cpp:105-111) */
    const int32_t* particle_layer;       /* [P] MP_LAYER_* */
    int64_t net_count;
    const int32_t* net_offsets;          /* [N+1] */
    const double* net_weights;           /* [N] */
    int64_t pin_count;
    const int32_t* pin_particles;        /* [I] particle index, or -1 for a fixed pin */
    const double* pin_offsets_or_static; /* [2I] oriented offset, or absolute position if fixed */
    int32_t bins_x, bins_y;
    double region_ll_x, region_ll_y, region_ur_x, region_ur_y;
    double target_density;
    const double* fixed_area;            /* [G] row-major, row = y */
    const double* dark_area;             /* [G] */
    int32_t device;                      /* HIP device ordinal */
    int32_t dct_precision;               /* MP_DCT_FP32 or MP_DCT_BF16X3 */
    uint64_t maximum_device_bytes;       /* allocation budget (GlobalPlacer.hpp:79) */
} mp_gp_static;

/* DensityMetrics, include/myplacement/metrics/Metrics.hpp:38-47 */
typedef struct mp_gp_density_metrics {
    double total_overflow_area;
    double normalized_overflow;
    double maximum_density;
    double average_density;
    double total_charge_area;
    double normalization_area;
    double placeable_area;
    double dark_area;
} mp_gp_density_metrics;

/* Scalars of CudaPlacementEvaluation, src/placement/CudaPlacementBackend.hpp:15-20 */
typedef struct mp_gp_scalars {
    double hpwl;
    double smooth_wirelength;
    double electrostatic_energy;
    mp_gp_density_metrics optimizer_density; /* with fillers */
    mp_gp_density_metrics design_density;    /* without fillers */
} mp_gp_scalars;

/* Adaptive-controller options, GlobalPlacer.hpp:26-80 (RUDY fields omitted: the RUDY term is
 * disabled by default and never runs on the device in the reference, AdaptiveGlobalPlacer.cpp:284). */
typedef struct mp_gp_options {
    int32_t iterations;                    /* 280 */
    int32_t bins_x, bins_y;                /* 64, 64 */
    double target_density;                 /* 0.85 */
    double maximum_bin_overflow;           /* 0.10 */
    double initial_smoothing;              /* 0.75 */
    double final_smoothing;                /* 0.12 */
    double maximum_movement_in_bins;       /* 0.22 */
    int32_t maximum_backtracks;            /* 6 */
    double backtracking_ratio;             /* 0.5 */
    double armijo_coefficient;             /* 1e-4 */
    double maximum_momentum;               /* 0.92 */
    double initial_movement_in_bins;       /* 0.12 */
    double objective_increase_for_density; /* 0.002 */
    double hpwl_growth_restart_threshold;  /* 0.02 */
    double density_stall_threshold;        /* 0.002 */
    double density_penalty_increase;       /* 1.05 */
    double density_penalty_decrease;       /* 0.88 */
    double minimum_density_penalty;        /* 1e-4 */
    double maximum_density_penalty;        /* 1e6 */
    double smoothing_overflow_exponent;    /* 1.0 */
    int32_t feasible_refinement_iterations;/* 12 */
    uint64_t maximum_fillers;              /* 50000 */
    uint64_t seed;                         /* 2026 */
    int32_t device;                        /* HIP device ordinal (0) */
    int32_t dct_precision;                 /* MP_DCT_* */
    uint64_t maximum_device_bytes;         /* budget for device allocations */
    double density_order_reuse_rows;       /* 2.5: the density deposit reuses its particle row order
                                              until positions may have moved this many bin rows;
                                              0 rebuilds it for every evaluation (results are
                                              identical either way; this is a speed setting) */
    int32_t capture_graphs;                /* 0: each iteration's launch chain is launched kernel by
                                              kernel (default); 1: it is captured into a HIP graph
                                              (one executable graph per chain shape, updated in place)
                                              and launched as one; results are identical */
    int32_t recompute_density;             /* 0: an iteration's evaluation at the candidate accepted by
                                              the iteration before takes that candidate's density,
                                              potential and energy, which are the same bits (default);
                                              1: it computes them again; results are identical */
} mp_gp_options;

/* GlobalPlacementIteration, GlobalPlacer.hpp:82-105 (same 22 columns as global_history.csv) */
typedef struct mp_gp_history_row {
    int32_t iteration;
    double hpwl, overflow, design_overflow, smooth_wirelength, density_energy, objective;
    double penalty, smoothing, step_size, maximum_displacement, gradient_norm, curvature;
    double rudy_energy, rudy_proxy_overflow, rudy_maximum_utilization, rudy_weight;
    int32_t backtracks;
    int32_t momentum_restarted, accepted, best_checkpoint, rudy_active;
} mp_gp_history_row;

/* GlobalPlacementResult, GlobalPlacer.hpp:107-129, plus evaluation counters. */
typedef struct mp_gp_result {
    double hpwl_before, hpwl_after, overflow_before, overflow_after;
    int32_t completed_iterations, accepted_iterations, momentum_restarts, rejected_candidates;
    int32_t best_checkpoint_iteration, restored_best_checkpoint;
    double best_checkpoint_hpwl, best_checkpoint_overflow;
    double elapsed_seconds;                 /* set-up + loop, as the reference reports (:656,951) */
    double loop_seconds;                    /* outer loop only */
    int64_t gradient_evaluations, value_evaluations;
    int64_t particle_count, filler_count;
    int32_t device_used;
    uint64_t reserved_device_bytes;
} mp_gp_result;

typedef struct mp_gp_backend mp_gp_backend;   /* one CudaPlacementBackend instance */
typedef struct mp_gp_session mp_gp_session;   /* device-resident adaptive GP run     */

/* ---- backend boundary --------------------------------------------------------------------- */

/* cudaPlacementBackendCompiled(), CudaPlacementBackend.hpp:44: 1 when the HIP kernels are built in. */
int mp_gp_compiled(void);

/* Thread-local text of the last failure (empty when none). */
const char* mp_gp_last_error(void);

/* tryCreateCudaPlacementBackend, CudaPlacementBackend.hpp:45-48 / .cu:1156-1216.
 * Uploads the static layout once.  MP_GP_UNAVAILABLE means a recoverable reason (no device,
 * budget exceeded, ...): the text is in mp_gp_last_error() and *out is NULL. */
mp_gp_status mp_gp_create(const mp_gp_static* layout, mp_gp_backend** out);

/* CudaPlacementBackend::evaluate, CudaPlacementBackend.hpp:32-33 / .cu:921-950.
 * xy: [2P] particle centres (FP64, particle order).  wire_grad / density_grad: [2P] outputs,
 * written only when want_grad != 0 (they may be NULL otherwise).  Blocking. */
mp_gp_status mp_gp_evaluate(mp_gp_backend* backend, const double* xy, double smoothing, int want_grad,
                            mp_gp_scalars* out, double* wire_grad, double* density_grad);

/* CudaPlacementBackend::device(), ::reservedBytes(), CudaPlacementBackend.hpp:34-35 */
int mp_gp_device(const mp_gp_backend* backend);
uint64_t mp_gp_reserved_bytes(const mp_gp_backend* backend);

/* ~CudaPlacementBackend() */
void mp_gp_destroy(mp_gp_backend* backend);

/* Copies of the device grids of the last evaluate (FP64 [G] each, NULL skips one): the centred
 * density source, potential (when computed), and the two field components. Test hook for the
 * field parity checks; reference has the same buffers in CudaNeumannField (.cu:625-762). */
mp_gp_status mp_gp_debug_grids(mp_gp_backend* backend, double* source, double* potential,
                               double* field_x, double* field_y);

/* Autotuner support (no counterpart in the reference, whose cuFFT transforms have no tile shapes):
 * the average time in microseconds of `reps` launches of one half-length DCT product of an
 * nx x ny grid (both multiples of 64) with tile shape `tile` (0..15, see gp_kernels.hpp), for
 * `evaluations` evaluations batched as the evaluation chain batches them, on device 0, with
 * operands of the chain's shapes.  myplacement_amd/autotune.py times every valid tile this way and
 * writes the table the library compiles in (csrc/dct_tiles.inc).  MP_GP_INVALID_ARGUMENT for a
 * tile that does not divide the product. */
mp_gp_status mp_gp_dct_tile_time(int32_t nx, int32_t ny, int32_t precision, int32_t mode, int32_t tile,
                                 int32_t evaluations, int32_t reps, double* microseconds);
/* One product of the same kind with tile `tile` on the caller's operands (no counterpart in the
 * reference; for tests of every tile): `data` holds the data grid of ny x 2 nx floats (two of them
 * for mode 3, the mirror-rows product), `factors` the 2 max(nx, ny)^2 factor floats, and `out`
 * receives as many floats as `data` holds.  The layouts are those of the evaluation chain's
 * products, mode 0 fold-cols, 1 fold-rows, 2 mirror-cols, 3 mirror-rows (gp_kernels.hpp). */
mp_gp_status mp_gp_dct_tile_run(int32_t nx, int32_t ny, int32_t precision, int32_t mode, int32_t tile,
                                const float* data, int64_t data_count, const float* factors, int64_t factor_count,
                                float* out);
/* The tile the library uses for a product (from its compiled table; 0 where the table names none). */
int32_t mp_gp_dct_tile_for(int32_t nx, int32_t ny, int32_t precision, int32_t mode);

/* ---- host layout services (GlobalPlacementSession set-up) --------------------------------- */

/* Default option values of GlobalPlacementOptions (GlobalPlacer.hpp:26-80). */
void mp_gp_default_options(mp_gp_options* options);

/* A prepared design: refreshDerivedData (PlacementDatabase.cpp:162-200), static density
 * (GlobalPlacementSupport.cpp:21-30), fillers (:52-103), HostLayout (.cu:784-839) and the
 * preconditioner base (AdaptiveGlobalPlacer.cpp:358-375). */
typedef struct mp_gp_prepared mp_gp_prepared;

mp_gp_status mp_gp_prepare(const mp_design* design, const mp_gp_options* options, mp_gp_prepared** out);
void mp_gp_prepared_destroy(mp_gp_prepared* prepared);

/* Sizes of a prepared design. */
typedef struct mp_gp_prepared_info {
    int64_t movable_count, filler_count, particle_count, net_count, pin_count, bin_count;
    double core_ll_x, core_ll_y, core_ur_x, core_ur_y;
    double nominal_row_height;
    double filler_width, filler_height;
} mp_gp_prepared_info;
mp_gp_status mp_gp_prepared_query(const mp_gp_prepared* prepared, mp_gp_prepared_info* info);

/* The flattened layout (pointers stay valid while `prepared` lives). */
mp_gp_status mp_gp_prepared_layout(const mp_gp_prepared* prepared, mp_gp_static* layout);

/* Movable module ids (movableModules() order) [movable_count]; filler centres [2F];
 * preconditioner base [P]. */
mp_gp_status mp_gp_prepared_arrays(const mp_gp_prepared* prepared, int64_t* movable_ids,
                                   double* filler_centers, double* preconditioner_base);

/* InitialPlacer randomPlacement, src/placement/InitialPlacer.cpp:102-120: mt19937_64(seed),
 * uniform centre per movable module, then clampModuleToCore.  centers: [2M] in/out. */
mp_gp_status mp_gp_random_placement(const mp_design* design, uint64_t seed, double* centers);

/* InitialPlacementOptions, quadratic fields (include/myplacement/placement/InitialPlacer.hpp:21-25). */
typedef struct mp_qp_options {
    int32_t outer_iterations;      /* 5 */
    int32_t solver_iterations;     /* 100: BiCGSTAB iterations per axis and outer iteration */
    double tolerance;              /* 1e-5: relative residual */
    double minimum_distance;       /* 1.0 */
    uint64_t clique_degree_limit;  /* 96: larger nets use the star model */
    int32_t device;                /* HIP device ordinal (0) */
} mp_qp_options;

/* InitialPlacementResult (InitialPlacer.hpp:28-34) for the quadratic method, plus solver figures. */
typedef struct mp_qp_result {
    double hpwl_before, hpwl_after;
    double elapsed_seconds;
    int32_t iterations;            /* completed outer iterations */
    int32_t solver_iterations;     /* BiCGSTAB iterations summed over axes and outer iterations */
    double relative_residual[2];   /* final |b - Ax| / |b| of the last outer iteration, x then y */
} mp_qp_result;

void mp_gp_default_qp_options(mp_qp_options* options);

/* InitialPlacer quadraticPlacement, src/placement/InitialPlacer.cpp:315-402: movable modules start
 * at the core centre; each outer iteration builds the clique (degree <= limit) / star bound-to-bound
 * system from the current pin positions with a 1e-4 anchor and solves x and y with unpreconditioned
 * BiCGSTAB on the device from the current positions, then clamps into the core.  centers: [2M]
 * in/out (fixed modules are read, movable ones written). */
mp_gp_status mp_gp_quadratic_placement(const mp_design* design, const mp_qp_options* options, double* centers,
                                       mp_qp_result* result);

/* ---- device-resident adaptive run (runAdaptiveGlobalPlacement with --compute-backend hip) ---- */

/* Creates the backend and the device-resident optimizer state from a prepared design.
 * centers: [2M] module centres (movable ones are read). */
mp_gp_status mp_gp_session_create(const mp_gp_prepared* prepared, const mp_gp_options* options,
                                  const double* centers, mp_gp_session** out);

/* Runs up to `iterations` more outer iterations of the adaptive loop (AdaptiveGlobalPlacer.cpp:735-928);
 * *completed receives how many ran; *stopped is 1 when the loop finished (budget or early stop). */
mp_gp_status mp_gp_session_step(mp_gp_session* session, int32_t iterations, int32_t* completed,
                                int32_t* stopped);

/* Restores the best checkpoint (AdaptiveGlobalPlacer.cpp:930-940) and writes module centres [2M]
 * (movable ones updated) and the result record. */
mp_gp_status mp_gp_session_finish(mp_gp_session* session, double* centers, mp_gp_result* result);

/* History rows so far; returns the row count when rows == NULL. */
int64_t mp_gp_session_history(const mp_gp_session* session, mp_gp_history_row* rows, int64_t capacity);

/* Current particle positions [2P] (host copy), in the prepared particle order. */
mp_gp_status mp_gp_session_positions(mp_gp_session* session, double* xy);

/* The session keeps its particles renumbered on the device (along a Z-order curve of the starting
 * positions, for locality).  order[k] = the prepared particle index of device particle k. */
mp_gp_status mp_gp_session_particle_order(const mp_gp_session* session, int64_t* order);

/* Per-kernel timing of the session's stream with HIP events (enable, then read).  enable: 0 off,
 * otherwise a bit mask of the kernel classes to time (bit i: the i-th name; -1: all).  The read
 * returns the class count; names: comma-separated kernel classes; per class the summed event time,
 * the launch count and the summed algorithmic work (bytes, or flops for the DCT products) of the
 * timed launches. */
mp_gp_status mp_gp_session_enable_kernel_timing(mp_gp_session* session, int enable);
/* The same, timing only every `every`-th launch of each enabled class (every >= 1), so that fewer
 * events enter the stream; the read then covers the timed launches only.  No counterpart in the
 * reference (measurement support). */
mp_gp_status mp_gp_session_enable_kernel_timing_every(mp_gp_session* session, int enable, int every);
int32_t mp_gp_session_kernel_timing(mp_gp_session* session, char* names, size_t names_len, double* total_us,
                                    int64_t* launches, double* work, int32_t capacity);

/* Access to the backend owned by a session (for evaluate-level checks); it takes and returns
 * particles in the device order of mp_gp_session_particle_order. */
mp_gp_backend* mp_gp_session_backend(mp_gp_session* session);

void mp_gp_session_destroy(mp_gp_session* session);

/* Whole GlobalPlacer::run (GlobalPlacer.cpp:134-146) for the adaptive optimizer on the HIP
 * backend: prepare, create, loop, restore.  centers: [2M] in/out. history may be NULL. */
mp_gp_status mp_gp_global_place(const mp_design* design, const mp_gp_options* options, double* centers,
                                 mp_gp_result* result, mp_gp_history_row* history, int64_t history_capacity);

/* mp_gp_global_place for `count` independent designs at once (no counterpart in the reference,
 * which places one design per process).  The runs advance in lockstep rounds: each round runs
 * every job's recorded launches of one stage (sorting, wirelength, density, each DCT product,
 * gather, the controller passes) as ONE batched launch over all jobs, and one reduction finalize
 * for all, then every job's controller decides its next step.  Each job's result is bit-equal to
 * its own mp_gp_global_place run.  All jobs run on options[0].device and hold their device state
 * together.  centers[i]: [2 M_i] in/out; results, statuses: [count] (may be NULL).  Returns the
 * first failing job's status (MP_GP_OK when all succeed); statuses[i] has each job's own.  From
 * the first batch or service on, the process keeps freed host memory in its heap (glibc trim
 * threshold and top pad raised), so later jobs' set-up arrays reuse touched pages. */
mp_gp_status mp_gp_global_place_batch(int32_t count, const mp_design* designs, const mp_gp_options* options,
                                       double* const* centers, mp_gp_result* results, mp_gp_status* statuses);

/* ---- placement service (no counterpart in the reference, which places one design per process) --
 * Jobs of mixed sizes are admitted against a device-memory budget into `max_active_jobs` slots.
 * MP_GP_SERVE_STREAMS: each runs mp_gp_global_place to completion on its own stream and worker
 * thread, so admitted jobs overlap on the device.  MP_GP_SERVE_BATCHED: one driver runs every
 * admitted job in the lockstep rounds of mp_gp_global_place_batch, and admits waiting jobs between
 * rounds (continuous batching).  A job's result does not depend on what runs beside it.  Creating
 * a service makes the process keep freed host memory in its heap (see mp_gp_global_place_batch). */
typedef enum mp_gp_serve_mode { MP_GP_SERVE_STREAMS = 0, MP_GP_SERVE_BATCHED = 1 } mp_gp_serve_mode;

typedef struct mp_gp_server_options {
    int32_t device;                /* HIP device ordinal (0) */
    int32_t max_active_jobs;       /* 8: jobs running at once */
    uint64_t memory_budget_bytes;  /* 200 GiB: admission budget for the jobs' device footprints */
    int32_t mode;                  /* mp_gp_serve_mode (MP_GP_SERVE_STREAMS) */
    int32_t reserved;
} mp_gp_server_options;

typedef struct mp_gp_job_report {
    double submit_seconds, start_seconds, finish_seconds;  /* since the service was created */
    double queue_seconds;          /* submit -> admitted */
    double latency_seconds;        /* submit -> result */
    uint64_t estimated_device_bytes;
    mp_gp_result result;
} mp_gp_job_report;

typedef struct mp_gp_server mp_gp_server;

void mp_gp_default_server_options(mp_gp_server_options* options);
/* Upper bound of the device bytes one mp_gp_global_place run of this design reserves. */
uint64_t mp_gp_estimate_device_bytes(const mp_design* design, const mp_gp_options* options);
mp_gp_status mp_gp_server_create(const mp_gp_server_options* options, mp_gp_server** out);
/* Copies the design; the caller may free it on return.  job: id for mp_gp_server_wait. */
mp_gp_status mp_gp_server_submit(mp_gp_server* server, const mp_design* design, const mp_gp_options* options,
                                 int64_t* job);
/* Blocks until the job is done; centers: [2M] placed centres (may be NULL).  Collects the job. */
mp_gp_status mp_gp_server_wait(mp_gp_server* server, int64_t job, double* centers, mp_gp_job_report* report);
/* Finishes the running jobs, drops the waiting ones and frees the service. */
void mp_gp_server_destroy(mp_gp_server* server);

/* ---- detailed placement: four-cell window reordering on the device -----------------------
 * Replaces tryRunCudaDetailedPlacement (src/placement/CudaDetailedPlacementBackend.hpp:14-17,
 * cuda/CudaDetailedPlacementBackend.cu:523-711) with DetailedPlacementOptions /
 * DetailedPlacementResult (include/myplacement/placement/DetailedPlacer.hpp:25-57) as C structs.
 * Runs on a legal placement: windows of four abutting standard cells are reordered in place. */
enum { MP_DP_NONE = 0, MP_DP_ADJACENT_SWAP = 1, MP_DP_WINDOW_REORDER = 2 };

typedef struct mp_dp_options {
    int32_t method;               /* MP_DP_WINDOW_REORDER (the only method the device pass runs) */
    int32_t passes;               /* 2: staggered window alignments */
    int32_t window_size;          /* 4 (the only size the device pass runs) */
    uint64_t maximum_net_degree;  /* 64: windows touching a larger net are left as they are */
    double improvement_epsilon;   /* 1e-9 */
    int32_t device;               /* HIP device ordinal (0) */
    uint64_t maximum_device_bytes;/* 2 GiB */
} mp_dp_options;

typedef struct mp_dp_result {
    double hpwl_before, hpwl_after;      /* calculateHpwl of the whole design */
    double elapsed_seconds;
    uint64_t evaluated_windows;
    uint64_t evaluated_permutations;
    uint64_t accepted_operations;
    int32_t completed_passes;
    int32_t device_used;                 /* -1 when no device pass ran (method none, no rows) */
    uint64_t reserved_device_bytes;      /* peak */
} mp_dp_result;

void mp_dp_default_options(mp_dp_options* options);
/* Starts from design->center / design->orientation; on success writes the placement to centers
 * [2M] and orientation [M] (may be NULL).  row_orientation [R]: orientation code of each row that
 * placed cells take (NULL: North, the SiteRow default; 8 = Unknown keeps the cell's own).
 * MP_GP_INVALID_ARGUMENT for options DetailedPlacer::run rejects (DetailedPlacer.cpp:50-71);
 * MP_GP_UNAVAILABLE (with mp_gp_last_error) for another method or window size, or no usable
 * device -- the caller then runs the CPU reference placer, as the reference does.  The output
 * arrays are untouched unless the call succeeds. */
mp_gp_status mp_dp_run(const mp_design* design, const int32_t* row_orientation, const mp_dp_options* options,
                       double* centers, int32_t* orientation, mp_dp_result* result);

/* mp_dp_run without any device allocation.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  Same arguments, results and statuses as mp_dp_run, bit for bit; every
 * device buffer is carved from a fixed scratch array the library holds on each GPU it runs on
 * (mp_dp_noalloc_scratch_bytes(), 64 MiB, kept until the process ends), and the work runs on the
 * null stream (no stream is created).  The design must fit that scratch: 56 bytes per module
 * + 24 per pin + 12 per net, each array rounded up to 256 bytes; a larger design gives
 * MP_GP_INVALID_ARGUMENT with the sizes in mp_gp_last_error() before anything runs.
 * maximum_device_bytes is not used.  Calls on one GPU take turns (they share its scratch); each
 * call overwrites what the previous one left there. */
mp_gp_status mp_dp_run_noalloc(const mp_design* design, const int32_t* row_orientation, const mp_dp_options* options,
                               double* centers, int32_t* orientation, mp_dp_result* result);
uint64_t mp_dp_noalloc_scratch_bytes(void);

#ifdef __cplusplus
}
#endif

#endif /* MP_GP_BACKEND_H */
