/* CPU FP64 restatement of the reference forward solve (SIMPLE + SST k-omega).
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker for the HIP engine:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * it. The product library never links, loads or calls it.
 *
 * All arrays are plain C arrays (structure of arrays). The mesh is a 2D
 * finite-volume mesh extruded one unit in z: every face normal has a zero z
 * component, so every z term of the reference's Vec3 arithmetic is an exact
 * zero and is omitted without changing any rounding.
 */
#ifndef RANS_ORACLE_H
#define RANS_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Boundary-condition kinds, in the order of the reference's BCType enum
 * (core/include/BoundaryCondition.hpp:13-23). */
enum { OR_BC_DIRICHLET = 0, OR_BC_NEUMANN = 1, OR_BC_INLET = 2, OR_BC_OUTLET = 3,
       OR_BC_WALL_NOSLIP = 4, OR_BC_WALL_KOMEGA = 5, OR_BC_WALL_MOVING = 6,
       OR_BC_SYMMETRY = 7, OR_BC_CYCLIC = 8 };
/* Patch kinds (the reference's Patch::type strings). */
enum { OR_PATCH_WALL = 0, OR_PATCH_INLET = 1, OR_PATCH_OUTLET = 2, OR_PATCH_SYMMETRY = 3,
       OR_PATCH_OTHER = 4 };
/* Linear solvers (LinearSolver.hpp:988-998) and convection schemes (FlowFields.hpp:144-145). */
enum { OR_SOLVER_PCG = 0, OR_SOLVER_BICGSTAB = 1, OR_SOLVER_AMG = 2 };
enum { OR_CONV_UPWIND = 0, OR_CONV_LINEAR_UPWIND = 1, OR_CONV_LIMITED = 2 };
/* Evaluation status (core/include/EvaluationTypes.hpp:6). */
enum { OR_CONVERGED = 0, OR_UNCONVERGED = 1, OR_DIVERGED = 2, OR_INVALID = 3, OR_UNKNOWN = 4 };
/* Observation kinds handled (ObservationOperator.hpp:33-40 subset). */
enum {
    OR_OBS_VELOCITY = 0,
    OR_OBS_PRESSURE = 1,
    OR_OBS_SKIN_FRICTION = 2,
    OR_OBS_DRAG = 3,
    OR_OBS_SEPARATION = 4,
    OR_OBS_REATTACHMENT = 5
};

typedef struct {
    int n_cells, n_internal, n_faces, n_patches;
    const int *owner, *neighbor;                   /* [n_faces], neighbor < 0 on boundary */
    const double *fcx, *fcy, *fnx, *fny, *farea;   /* face centre, unit normal, area */
    const double *fdelta, *fdx, *fdy, *fweight;    /* |d|, d, interpolation weight */
    const double *fshx, *fshy;                     /* periodic shift of the neighbour */
    const double *ccx, *ccy, *vol, *wall_dist;     /* [n_cells] */
    const int *cell_face_ptr, *cell_face_idx;      /* per-cell face lists, reference order */
    const int *patch_type;                         /* [n_patches] OR_PATCH_* */
    const int *patch_face_ptr, *patch_face_idx;    /* per-patch face lists */
    /* front and back faces of a one-layer extruded mesh folded out of the
     * face arrays: per cell (area, delta) twice in face order, area 0 where
     * absent; NULL for a planar mesh (core/src/Mesh.cpp:119-320) */
    const double *cell_folded;
} or_mesh_t;

typedef struct {
    const int *u_type, *p_type, *k_type, *w_type; /* [n_patches] OR_BC_* */
    const double *u_val;                          /* [2*n_patches] (x, y) */
    const double *p_val, *k_val, *w_val;          /* [n_patches] */
} or_bc_t;

typedef struct {
    int max_iterations;
    double convergence_tol, divergence_limit;
    int convergence_window, monitor_interval;
    double alpha_u, alpha_p, alpha_k, alpha_omega;
    int inner_iterations;
    double inner_tolerance;
    int turb_start_iter, turb_update_interval, nut_floor_iters;
    int momentum_convection, turbulence_convection;
    double body_force_x, body_force_y;
    double k_min, k_max, omega_min;
    int pressure_solver, momentum_solver, turbulence_solver;
    int sst_variant; /* 0 Full, 1 NoLimiter, 2 KOmega */
} or_settings_t;

typedef struct {
    int n_obs;
    const int *type, *component;
    const double *observed, *variance;                 /* total variance per observation */
    const double *ref_velocity, *ref_pressure, *ref_density;
    /* probe stencils (velocity and pressure probes; -1 cell otherwise) */
    const int *cell;
    const double *offx, *offy, *axx, *axy, *ayy, *det;
    const int *nb_ptr, *nb_cell, *nb_face;             /* nb_cell < 0: boundary face value */
    const double *nb_dx, *nb_dy, *nb_w;
    /* skin-friction stencils */
    const int *face_a, *face_b;
    const double *t;
} or_obs_t;

typedef struct {
    int status, iters, stalled, converged, diverged;
    double worst, rate, loglik;
    double *predictions;                       /* [n_obs], caller-owned */
    /* optional field outputs, caller-owned, may be NULL */
    double *Ux, *Uy, *p, *k, *omega, *nuT, *F1, *F2, *Pk, *CDkw; /* [n_cells] */
    double *flux;                                                /* [n_faces] */
    long long krylov_iters[5];                 /* Ux, Uy, p', k, omega inner iterations */
    char error[256];
} or_result_t;

/* One forward evaluation: theta is already unpacked into the 11 SST
 * coefficients (InferenceParameters.hpp:8-12 order). Returns 0, or nonzero
 * with result->error set for an invalid problem (the result then carries
 * status OR_UNKNOWN). Follows ForwardModel::evaluateWithDiagnostics
 * (rans-uq/src/ForwardModel.cpp:35-137) without the warm-start cache. */
int or_evaluate(const or_mesh_t *mesh, const or_bc_t *bc, const or_settings_t *settings,
                double nu, const double coeffs[11], const double init[5], const or_obs_t *obs,
                or_result_t *result);

/* or_evaluate with a warm start and options. warm (NULL: cold start) and
 * save (NULL: not written) hold a flow state: the interior fields Ux, Uy, p,
 * k, omega, nuT, F1, F2, Pk, CDkw (n_cells each), the boundary values Ux, Uy,
 * p, k, omega (max(n_boundary, 1) each), the turbulence-established flag and
 * one pad. flags: OR_FLAG_NO_MONITOR drops the observation convergence
 * monitor, OR_FLAG_WARM_BCS re-applies the boundary conditions to the warm
 * state with this solve's beta1 (ParameterSensitivity.cpp:790-794).
 * tol_override / max_iter_override > 0 replace the settings' values. */
enum { OR_FLAG_NO_MONITOR = 1, OR_FLAG_WARM_BCS = 2 };
int or_evaluate_ex(const or_mesh_t *mesh, const or_bc_t *bc, const or_settings_t *settings,
                   double nu, const double coeffs[11], const double init[5], const or_obs_t *obs,
                   const double *warm, double *save, int flags, double tol_override,
                   int max_iter_override, or_result_t *result);

/* Closed-form pointwise SST functions, exported for known-answer tests
 * (core/src/SSTModel.cpp:15-85,102-108). c[11] is the coefficient vector. */
double or_sst_f1(const double c[11], double k, double omega, double y, double nu, double cdkw);
double or_sst_f2(const double c[11], double k, double omega, double y, double nu);
double or_sst_nut(const double c[11], int variant, double k, double omega, double s, double f2);
double or_sst_pk(const double c[11], double nut, double s, double k, double omega);
double or_sst_pomega(const double c[11], double nut, double s, double k, double omega);

/* Linear-solver entry points on an LDU system, exported so the Krylov and
 * AMG restatements can be checked against exact solutions. */
/* Experiment switch (off by default, process-wide): round every assembled
 * system to FP32 before it is used; residuals, norms and solves stay FP64. */
void or_set_assembly_rounding(int on);

int or_linear_solve(int solver, int n, int nif, const int *own, const int *nbr, const double *diag,
                    const double *upper, const double *lower, const double *source, double *x,
                    int max_iter, double tol, int *iterations, double *initial_res,
                    double *final_res);

#ifdef __cplusplus
}
#endif
#endif
