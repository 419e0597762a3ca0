/* C-ABI of the MI355X forward-solve engine (SIMPLE + SST k-omega).
 *
 * This is the drop-in boundary for the reference's batch forward model:
 *   ForwardBatch(mesh, params, observations, bcs, nu, settings, u, p, k, omega)
 *     rans-uq/include/ForwardBatch.hpp:34-50  -> rq_engine_create
 *   ForwardBatch::evaluate(theta, workers)
 *     rans-uq/include/ForwardBatch.hpp:52-124 -> rq_engine_evaluate
 *   (its Python binding: rans-uq/python/bindings.cpp:888-916)
 * Plain pointers and sizes only; every error is a return code plus text in
 * the caller's buffer (no exceptions cross this boundary).
 *
 * Status codes follow the reference's EvaluationStatus order
 * (core/include/EvaluationTypes.hpp:6): 0 Converged, 1 Unconverged,
 * 2 Diverged, 3 InvalidParameters, 4 Unknown.
 */
#ifndef RANS_AMD_H
#define RANS_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RQ_OK 0
#define RQ_EINVAL 1   /* invalid argument (std::invalid_argument in the reference) */
#define RQ_EDEVICE 2  /* HIP runtime failure or no device */
#define RQ_ENOMEM 3   /* allocation failure (std::bad_alloc, fatal for a batch) */

/* A 2D finite-volume mesh (all face normals in the x-y plane), internal faces
 * first: the flattened Mesh of core/include/Mesh.hpp:52-268 with its wall
 * distance (Mesh::computeWallDistance, core/src/Mesh.cpp:431-463). */
typedef struct rq_mesh_desc {
    int32_t n_cells, n_internal, n_faces, n_patches;
    const int32_t *owner, *neighbor;         /* [n_faces], neighbor -1 on boundary */
    const double *face_center, *face_normal; /* [3*n_faces] */
    const double *face_area, *face_delta;    /* [n_faces] */
    const double *face_d;                    /* [3*n_faces] owner -> neighbour centre */
    const double *face_weight;               /* [n_faces] */
    const double *face_shift;                /* [3*n_faces] periodic shift */
    const double *cell_center;               /* [3*n_cells] */
    const double *cell_volume, *wall_distance;
    const int32_t *cell_face_ptr, *cell_face_idx; /* each cell's faces, ascending */
    const int32_t *patch_type;               /* 0 wall 1 inlet 2 outlet 3 symmetry 4 other */
    const int32_t *patch_face_ptr, *patch_face_idx;
    /* Front and back faces of a one-layer extruded mesh (an OpenFOAM 2D
     * case, core/src/Mesh.cpp:119-320), folded out of the face arrays:
     * per cell (area, delta) of up to two such faces in face order, area 0
     * where a cell has fewer; NULL for a planar mesh. Each adds a
     * zero-gradient boundary term (no flux) to the cell's momentum, k and
     * omega rows (SIMPLESolver.cpp:360-382, 983-1001, 1092-1113). */
    const double *cell_folded; /* [4*n_cells] or NULL */
} rq_mesh_desc;

/* FlowBoundaryConditions (core/include/BoundaryCondition.hpp:105-167):
 * per-patch BCType codes (BoundaryCondition.hpp:13-23 order) and uniform
 * values. */
typedef struct rq_bc_desc {
    const int32_t *u_type, *p_type, *k_type, *w_type; /* [n_patches] */
    const double *u_val;                              /* [2*n_patches] */
    const double *p_val, *k_val, *w_val;              /* [n_patches] */
} rq_bc_desc;

/* SolverSettings (core/include/FlowFields.hpp:70-241). Solvers: 0 PCG,
 * 1 BiCGSTAB, 2 AMG. Schemes: 0 upwind, 1 linearUpwind, 2 limitedLinearUpwind. */
typedef struct rq_settings {
    int32_t max_iterations;
    double convergence_tol, divergence_limit;
    int32_t convergence_window, monitor_interval;
    double alpha_u, alpha_p, alpha_k, alpha_omega;
    int32_t inner_iterations;
    double inner_tolerance;
    int32_t turb_start_iter, turb_update_interval, nut_floor_iters;
    int32_t momentum_convection, turbulence_convection;
    double body_force_x, body_force_y;
    double k_min, k_max, omega_min;
    int32_t pressure_solver, momentum_solver, turbulence_solver;
    int32_t sst_variant;
    /* engine options with no reference counterpart */
    int32_t mixed_precision; /* 1: FP32 Krylov vectors, FP64 defect correction */
    int32_t block_size;      /* threads per draw; 0 = tuned default */
    int32_t profile;         /* 1: per-phase clock counts in rq_batch_out.phase_cycles */
    int32_t grid_blocks;     /* workgroups per draw: 0 = automatic (one, except arenas in HBM of
                                >= 4096 cells with no explicit block_size, which spread over
                                the CUs the batch leaves idle, one per 256 cells at most);
                                1 = one; 2..64 = that many co-resident workgroups */
    int32_t amg_block_rows;  /* multi-workgroup draws: multigrid levels of at most this many
                                rows run on one workgroup of the draw, larger ones on all of
                                them (0 = 1024); the results do not depend on it */
    int32_t grid_plain_launch; /* multi-workgroup draws: 0 = cooperative launch (co-residency
                                  guaranteed by the runtime); 1 = ordinary dispatch of the same
                                  grid, sized by the same occupancy check, for tools that cannot
                                  trace cooperative launches. Every grid-barrier wait is bounded,
                                  so a grid that is not co-resident ends as Unknown rows. */
} rq_settings;

/* InferenceParameterSet (core/include/InferenceParameters.hpp:18-188). */
typedef struct rq_param_desc {
    int32_t n_active;
    const int32_t *active_indices; /* into the 11 SST coefficients */
    const double *base;            /* [11] values of the inactive coefficients */
    const double *lower, *upper;   /* [n_active] bounds */
} rq_param_desc;

/* PreparedObservation stencils (core/include/ObservationOperator.hpp:611-669):
 * type 0 velocity probe, 1 pressure tap, 2 skin friction, 3 drag, 4 separation
 * point, 5 reattachment length. For types 3-5 the stencil
 * nb_ptr[i]..nb_ptr[i+1] lists the patch's faces in patch order (nb_cell -1,
 * nb_face the boundary face, nb_dx its centre x); drag takes its reference
 * area in offx[i]. */
typedef struct rq_obs_desc {
    int32_t n_obs;
    const int32_t *type, *component;
    const double *observed, *variance;
    const double *ref_velocity, *ref_pressure, *ref_density;
    const int32_t *cell;
    const double *offx, *offy, *axx, *axy, *ayy, *det;
    const int32_t *nb_ptr, *nb_cell, *nb_face;
    const double *nb_dx, *nb_dy, *nb_w;
    const int32_t *face_a, *face_b;
    const double *t;
} rq_obs_desc;

typedef struct rq_engine rq_engine;

/* Per-row outputs of one batch (ForwardBatchRow, ForwardBatch.hpp:8-19).
 * Any pointer except status may be NULL. predictions is n_rows x n_obs, NaN
 * for rows without predictions. row_errors is n_rows x row_error_stride
 * bytes of NUL-terminated text. fields, if given, receives per row 10
 * n_cells blocks (Ux Uy p k omega nuT F1 F2 Pk CDkw) then n_faces face fluxes. */
typedef struct rq_batch_out {
    int32_t *status;
    double *loglik, *predictions;
    int32_t *iterations;
    double *final_worst_measure, *contraction_rate_u;
    uint8_t *stalled, *converged;
    char *row_errors;
    size_t row_error_stride;
    double *fields;
    int64_t *krylov_iterations; /* n_rows x 5: Ux, Uy, p', k, omega */
    int64_t *phase_cycles;      /* n_rows x 16 shader clocks per solve phase (profile = 1);
                                   multi-workgroup draws: [14] clocks rank 0 waited in grid
                                   barriers, [15] the number of grid barriers */
} rq_batch_out;

/* ForwardBatch constructor: copies every input (the caller may free or
 * change its arrays afterwards). init = {Ux, Uy, p, k, omega}. */
rq_engine *rq_engine_create(const rq_mesh_desc *mesh, const rq_param_desc *params,
                            const rq_obs_desc *obs, const rq_bc_desc *bcs, double nu,
                            const rq_settings *settings, const double init[5], char *err,
                            size_t errcap);

/* ForwardBatch::evaluate: theta is n_rows x n_params (row-major). All rows
 * are solved on the GPU together; workers must be >= 1 and only keeps the
 * reference's argument contract. Like the reference's const evaluate, it may
 * be called from several threads on one engine: an internal lock runs the
 * calls (and autotune, graph-replay and timing calls) one after another. */
int rq_engine_evaluate(rq_engine *engine, const double *theta, int64_t n_rows, int32_t n_params,
                       int32_t workers, rq_batch_out *out, char *err, size_t errcap);

/* No counterpart in the reference: rq_engine_evaluate for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). The same kernels, with every per-call
 * device buffer (slot table, result rows, predictions, grid-barrier state,
 * HBM arenas, multigrid workspace) taken from one fixed device region of
 * RQ_NOALLOC_SCRATCH_BYTES per GPU, reserved when the library's kernels are
 * loaded on that GPU and held until the process ends; the call itself makes
 * no hipMalloc, hipFree or other allocation. A call that would need more
 * than the region (about rows x (arena + multigrid workspace + result and
 * prediction bytes)) fails with RQ_EINVAL and names the size. out->fields
 * must be NULL and, when the problem has observations, out->predictions
 * must be given. The region keeps the last call's contents; every call on
 * one GPU shares it, so calls that overlap on a GPU would race on it: the
 * library runs them one at a time with a process-wide lock per GPU. No graph
 * replay: the launch is direct. */
#define RQ_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
int rq_engine_evaluate_noalloc(rq_engine *engine, const double *theta, int64_t n_rows,
                               int32_t n_params, int32_t workers, rq_batch_out *out, char *err,
                               size_t errcap);

int32_t rq_engine_n_observations(const rq_engine *engine);
int32_t rq_engine_n_parameters(const rq_engine *engine);
int32_t rq_engine_n_cells(const rq_engine *engine);
int32_t rq_engine_n_faces(const rq_engine *engine);

/* Device time of the last evaluate's solve kernel in milliseconds (HIP
 * events on the engine's stream) and its launch shape (workgroup size,
 * arena placement and size, workgroups per draw). */
int rq_engine_last_timing(const rq_engine *engine, double *kernel_ms, int32_t *block_size,
                          int32_t *arena_in_lds, int64_t *arena_bytes, int32_t *grid_blocks);

void rq_engine_destroy(rq_engine *engine);

/* No counterpart in the reference: the launch-shape autotuner. Every
 * candidate workgroup size (64, 128, 256, 512)
 * and, for arenas in HBM of at least 1024 cells, every workgroups-per-draw
 * count 2..64 that the device can hold runs the given rows with each solve
 * capped at `iterations` outer iterations, in the order listed. The kept
 * shape is the first one in that order whose kernel time is within 3 % of
 * the fastest (shapes within timing noise of each other resolve the same way
 * on every device), and it becomes the engine's launch shape for later
 * evaluations. When the kept shape spreads draws with an AMG pressure solve
 * over several workgroups, the one-workgroup multigrid threshold
 * (amg_block_rows: the configured one, then 256, 512, 2048) is chosen the
 * same way at that shape. kernel_ms is NaN for a shape that could not run. */
typedef struct rq_tune_out {
    int32_t n_candidates;
    int32_t block_size[16], grid_blocks[16], amg_block_rows[16];
    double kernel_ms[16];
    int32_t chosen_block_size, chosen_grid_blocks, chosen_amg_block_rows;
} rq_tune_out;
int rq_engine_autotune(rq_engine *engine, const double *theta, int64_t n_rows, int32_t n_params,
                       int32_t iterations, rq_tune_out *out, char *err, size_t errcap);

/* No counterpart in the reference: HIP graph replay of the batch launch.
 * With enable = 1, an evaluation whose draws run one workgroup each captures
 * its launch (slot upload, solve kernel, result download through pinned host
 * buffers) into a HIP graph the first time and replays that graph while the
 * row count, launch shape, settings and buffers stay the same; draws spread
 * over several workgroups (cooperative launches) are launched directly.
 * kernel_ms of rq_engine_last_timing then covers the whole replayed graph. */
int rq_engine_set_graph_replay(rq_engine *engine, int32_t enable, char *err, size_t errcap);
/* Graph launches so far (each capture's first launch included). */
int64_t rq_engine_graph_launches(const rq_engine *engine);

/* No counterpart in the reference: the pressure multigrid hierarchy (for
 * byte models). Returns the number of levels (0 without AMG, -1 for a null
 * engine) and fills rows and CSR entries (coarse rows padded to multiples
 * of 4) of the first cap levels. */
int32_t rq_engine_amg_levels(const rq_engine *engine, int32_t *rows, int64_t *entries,
                             int32_t cap);

/* Library identity: compiler, target and build flags. */
const char *rq_build_info(void);

/* ---------------------------------------------------------------------------
 * Serving. No counterpart in the reference (its callers build a ForwardBatch
 * per batch); these entry points are the continuous-batching front of the
 * same solve: single-row requests on registered problems are admitted into
 * GPU slots as slots free up, and overlap on the device.
 * ------------------------------------------------------------------------- */
#define RQ_EAGAIN 4   /* request ring full: poll finished tickets, then resubmit */

typedef struct rq_server rq_server;

typedef struct rq_serve_options {
    int32_t max_slots;          /* concurrent solves at most (also capped by CU count) */
    int64_t hbm_budget_bytes;   /* arena memory for the slots; 0 = no limit */
    int32_t ring_size;          /* requests in flight at most */
    double idle_exit_seconds;   /* the device loop returns after this long without work */
} rq_serve_options;

/* One finished request (the ForwardBatchRow fields of one row). */
typedef struct rq_serve_row {
    int32_t status;
    double loglik;
    double *predictions;        /* caller-owned, n_observations of the problem; may be NULL */
    int32_t iterations;
    double final_worst_measure, contraction_rate_u;
    uint8_t stalled, converged;
    int64_t krylov_iterations[5];
    char error[256];
} rq_serve_row;

rq_server *rq_server_create(const rq_serve_options *options, char *err, size_t errcap);

/* Registers a problem (the ForwardBatch constructor's inputs); returns its id
 * or -1. All problems are registered before the first submit. */
int32_t rq_server_add_problem(rq_server *server, const rq_mesh_desc *mesh,
                              const rq_param_desc *params, const rq_obs_desc *obs,
                              const rq_bc_desc *bcs, double nu, const rq_settings *settings,
                              const double init[5], char *err, size_t errcap);

/* Submits one coefficient row; returns a ticket >= 0, -2 when the ring is
 * full, -1 on error. Rows out of bounds settle at once as InvalidParameters. */
int64_t rq_server_submit(rq_server *server, int32_t problem, const double *theta,
                         int32_t n_params, char *err, size_t errcap);

/* Non-blocking: *done = 1 and row filled once the ticket's solve finished. */
int rq_server_poll(rq_server *server, int64_t ticket, rq_serve_row *row, int32_t *done,
                   char *err, size_t errcap);

int rq_server_stats(const rq_server *server, int32_t *slots, int64_t *slot_arena_bytes,
                    int64_t *submitted, int64_t *completed);

void rq_server_destroy(rq_server *server);

/* ---------------------------------------------------------------------------
 * Single-row model with a warm-start cache: ForwardModel
 * (rans-uq/include/ForwardModel.hpp:191-247, rans-uq/src/ForwardModel.cpp:35-137)
 * and its WarmStartCache (ForwardModel.hpp:21-157). Converged flow states stay
 * in device memory; a row whose coefficients lie within the threshold of a
 * cached row starts from the nearest cached state instead of the uniform
 * initialisation.
 * ------------------------------------------------------------------------- */
typedef struct rq_model rq_model;

/* WarmStartCache(maxSize, threshold, byteBudget) (ForwardModel.hpp:33-37);
 * ForwardModel's defaults are 50, 0.5 and 64 MiB (ForwardModel.cpp:17). */
typedef struct rq_cache_options {
    int32_t max_size;
    double threshold;
    int64_t byte_budget;
} rq_cache_options;

/* ForwardModel constructor; cache NULL = the defaults. */
rq_model *rq_model_create(const rq_mesh_desc *mesh, const rq_param_desc *params,
                          const rq_obs_desc *obs, const rq_bc_desc *bcs, double nu,
                          const rq_settings *settings, const double init[5],
                          const rq_cache_options *cache, char *err, size_t errcap);

/* ForwardModel::evaluateWithDiagnostics on each row in turn (each row sees
 * the states the rows before it stored); used_warm_start (may be NULL) gets
 * ForwardEvaluationResult::usedWarmStart per row. out as for
 * rq_engine_evaluate. */
int rq_model_evaluate(rq_model *model, const double *theta, int64_t n_rows, int32_t n_params,
                      rq_batch_out *out, uint8_t *used_warm_start, char *err, size_t errcap);

/* No counterpart in the reference: every row looks up the cache as it stands
 * at the call, all rows run in one launch, and the converged ones are stored
 * afterwards in row order (for callers that evaluate a stencil of
 * neighbouring points at once, such as a pattern-search poll). */
int rq_model_evaluate_batch(rq_model *model, const double *theta, int64_t n_rows,
                            int32_t n_params, rq_batch_out *out, uint8_t *used_warm_start,
                            char *err, size_t errcap);

/* The model's engine, for rq_engine_n_* and rq_engine_last_timing. */
const rq_engine *rq_model_engine(const rq_model *model);

/* WarmStartCache::size / retainedBytes / byteBudget / setByteBudget / clear
 * (ForwardModel.hpp:93-116). */
int32_t rq_model_cache_size(const rq_model *model);
int64_t rq_model_cache_retained_bytes(const rq_model *model);
int64_t rq_model_cache_byte_budget(const rq_model *model);
int rq_model_cache_set_byte_budget(rq_model *model, int64_t bytes);
int rq_model_cache_clear(rq_model *model);
void rq_model_destroy(rq_model *model);

/* The cache's bookkeeping on its own, without a GPU: entries carry an integer
 * tag in place of a flow state. Same rules as rq_model's cache
 * (WarmStartCache::store / nearest / setByteBudget / clear / entryBytes). */
typedef struct rq_warm_cache rq_warm_cache;
rq_warm_cache *rq_warm_cache_create(int32_t max_size, double threshold, int64_t byte_budget,
                                    char *err, size_t errcap);
/* 1 stored, 0 not retained (non-finite row, zero capacity or over budget). */
int32_t rq_warm_cache_store(rq_warm_cache *cache, const double *theta, int32_t n, int64_t tag,
                            int64_t bytes);
/* Tag of the nearest entry within the threshold, or -1. */
int64_t rq_warm_cache_nearest(const rq_warm_cache *cache, const double *theta, int32_t n);
int32_t rq_warm_cache_size(const rq_warm_cache *cache);
int64_t rq_warm_cache_retained_bytes(const rq_warm_cache *cache);
int rq_warm_cache_set_byte_budget(rq_warm_cache *cache, int64_t bytes);
int rq_warm_cache_clear(rq_warm_cache *cache);
int64_t rq_warm_cache_entry_bytes(int32_t n_theta, int32_t n_cells, int32_t n_boundary);
void rq_warm_cache_destroy(rq_warm_cache *cache);

/* ---------------------------------------------------------------------------
 * Total sensitivities by warm-started central differences: ParameterSensitivity
 * (rans-uq/include/ParameterSensitivity.hpp:68-133) with solveState and
 * etaJacobianWarmFD (rans-uq/src/ParameterSensitivity.cpp:35-59,729-840).
 * Coefficient vectors are the eleven SST coefficients. The converged base
 * state stays in device memory; the 22 perturbed solves of a Jacobian start
 * from it (boundary conditions re-applied with each row's beta1, no
 * observation convergence monitor, each coefficient's own tolerance) and run
 * in one launch.
 * ------------------------------------------------------------------------- */
typedef struct rq_sensitivity rq_sensitivity;

/* TangentGradientResult of the warm-FD path (ParameterSensitivity.hpp:24-36).
 * Arrays are caller-owned; any may be NULL. Failed columns hold NaN. */
typedef struct rq_warm_fd_out {
    double *d_obs_d_theta;      /* n_obs x 11 (row-major): dObsDTheta */
    double *log_lik_gradient;   /* 11 */
    int32_t *iterations;        /* 11: SIMPLE iterations of the + and - solves together */
    double *worst_measure;      /* 11: the larger final convergence measure (inf if none) */
    int32_t *converged;         /* 11: both solves converged */
    double *steps;              /* 11: central half-widths */
    double *tolerances;         /* 11: perturbed-solve tolerances */
    int32_t *positive_status;   /* 11 */
    int32_t *negative_status;   /* 11 */
    int64_t *n_residual_evals;  /* 1: SIMPLE iterations of every solved row */
} rq_warm_fd_out;

/* ParameterSensitivity constructor (no parameter set: all eleven coefficients). */
rq_sensitivity *rq_sensitivity_create(const rq_mesh_desc *mesh, const rq_obs_desc *obs,
                                      const rq_bc_desc *bcs, double nu,
                                      const rq_settings *settings, const double init[5],
                                      char *err, size_t errcap);

/* solveState: a cold solve without the observation monitor; *status gets the
 * EvaluationStatus code, and a converged state becomes the base. */
int rq_sensitivity_solve_state(rq_sensitivity *sens, const double theta11[11], int32_t *status,
                               char *err, size_t errcap);

int32_t rq_sensitivity_has_state(const rq_sensitivity *sens);

/* etaJacobianWarmFD(theta11, hRel, hFloor, warmMaxIter, warmTol, absoluteSteps);
 * theta11 must be the base state's coefficients. absolute_steps: NULL or 11
 * positive half-widths. RQ_EINVAL for the reference's std::invalid_argument
 * and logic_error cases. */
int rq_sensitivity_eta_jacobian_warm_fd(rq_sensitivity *sens, const double theta11[11],
                                        double h_rel, double h_floor, int32_t warm_max_iter,
                                        double warm_tol, const double *absolute_steps,
                                        rq_warm_fd_out *out, char *err, size_t errcap);

/* The sensitivity's engine, for rq_engine_n_observations and timing. */
const rq_engine *rq_sensitivity_engine(const rq_sensitivity *sens);
void rq_sensitivity_destroy(rq_sensitivity *sens);

#ifdef __cplusplus
}
#endif
#endif
