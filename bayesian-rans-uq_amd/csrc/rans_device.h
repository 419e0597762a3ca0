// Plain-data structures shared by the host engine and the gfx950 kernels.
// Everything here is POD so it can be copied to the device as a kernel
// argument or into device memory.
#pragma once
#include <stdint.h>

namespace rq {

constexpr int kSlots = 4;          // faces per cell (ELL width): quadrilaterals and triangles
constexpr int kMaxWindow = 256;    // largest contraction window supported
constexpr int kPhases = 16;        // profiled phases of one outer iteration

enum BcType { BC_DIRICHLET = 0, BC_NEUMANN, BC_INLET, BC_OUTLET, BC_WALL_NOSLIP,
              BC_WALL_KOMEGA, BC_WALL_MOVING, BC_SYMMETRY, BC_CYCLIC };
enum PatchType { PATCH_WALL = 0, PATCH_INLET, PATCH_OUTLET, PATCH_SYMMETRY, PATCH_OTHER };
enum SolverKind { SOLVER_PCG = 0, SOLVER_BICGSTAB = 1, SOLVER_AMG = 2 };
enum Scheme { CONV_UPWIND = 0, CONV_LINEAR_UPWIND = 1, CONV_LIMITED = 2 };
enum Status { ST_CONVERGED = 0, ST_UNCONVERGED, ST_DIVERGED, ST_INVALID, ST_UNKNOWN };
enum ObsKind {
    OBS_VELOCITY = 0,
    OBS_PRESSURE = 1,
    OBS_SKIN_FRICTION = 2,
    OBS_DRAG = 3,          // pressure plus viscous x force on a patch, as Cd
    OBS_SEPARATION = 4,    // x where the wall shear first turns non-positive
    OBS_REATTACHMENT = 5,  // x where the wall shear first turns non-negative
};
// device error codes (mapped to the reference's exception texts on the host)
enum DevError { ERR_NONE = 0, ERR_NONFINITE_SYSTEM = 1, ERR_NONFINITE_MONITOR = 2,
                ERR_GRID_TIMEOUT = 3 };

// Multi-workgroup draws: the workgroups of one draw meet at grid barriers on
// this per-draw block (device memory, zeroed before every launch) and
// exchange their reduction partials through it.
// Fixed device region of the allocation-free evaluation
// (rq_engine_evaluate_noalloc): one per GPU, held until the process ends.
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
constexpr int kMaxDevices = 64;
constexpr int kMaxGridBlocks = 64;
constexpr int kGridBlock = 256;      // threads per workgroup of a multi-workgroup draw
constexpr int kXcds = 8;             // accelerator complex dies of an MI355X
struct GridState {
    unsigned long long count;  // barrier arrivals since launch (never reset)
    int abort, pad;            // 1: a barrier wait timed out; every workgroup leaves
    int xcc[kMaxGridBlocks];   // XCD of each workgroup of the draw
    double partials[2][kMaxGridBlocks][4];
};

// Per-draw arena, array-of-structures: every cell, face and boundary face is
// one record of doubles with an odd stride (so consecutive records start in
// different LDS banks), and a variable is a fixed offset inside its record.
// Geometry lives in the same records, pre-filled by the host.
enum CellVar {
    V_Ux, V_Uy, V_p, V_k, V_w, V_nuT, V_F1, V_F2, V_Pk, V_CDkw, V_S, V_aPs, V_UxP, V_UyP, V_pP,
    V_gUxx, V_gUxy, V_gUyx, V_gUyy, V_gPx, V_gPy, V_gKx, V_gKy, V_gWx, V_gWy,
    V_diag, V_src, V_off0, V_off1, V_off2, V_off3, V_wallv,
    V_kx, V_kr, V_krh, V_kp, V_kv, V_ks, V_kt, V_kph, V_ksh,
    V_vol, V_oms, V_oml, V_wd, V_COUNT
};
// face record: transported state, then geometry Sx=nx*A, Sy=ny*A, A,
// dlt=max(delta,1e-20), w, wm=1-w, nx, ny, ex/ey=d/max(delta,1e-30), the
// upwind offset vectors for owner-upwind (dO) and neighbour-upwind (dN)
// faces, and the planar-balance wall-normal gradient omg.
enum FaceVar {
    W_phi, W_phiBar, W_phiPred, W_pCoeff, W_Sx, W_Sy, W_A, W_dlt, W_w, W_wm, W_nx, W_ny, W_ex,
    W_ey, W_dOx, W_dOy, W_dNx, W_dNy, W_omg, W_COUNT
};
// boundary-face record: face values, then the prescribed values of the
// face's patch (velocity, pressure, k, omega)
enum BndVar { B_Ux, B_Uy, B_p, B_k, B_w, B_pP, B_uvx, B_uvy, B_pv, B_kv, B_wv, B_COUNT };
// observation record: probe stencil (offset, 2x2 least-squares system),
// skin-friction blend, references and the monitor floor
enum ObsVar { O_offx, O_offy, O_axx, O_axy, O_ayy, O_det, O_t, O_refv, O_refp, O_refrho, O_floor,
              O_COUNT };
constexpr int kNA = V_COUNT | 1;
constexpr int kNF = W_COUNT | 1;
constexpr int kNB = B_COUNT | 1;
constexpr int kNO = O_COUNT | 1;
// boundary flag word: bit 0 wall, bit 1 outlet, then the patch's BC types
// for velocity, pressure, k and omega in 4-bit fields from bit 4
constexpr int kFlagWall = 1, kFlagOutlet = 2;
constexpr int bc_type(int flag, int field) { return (flag >> (4 + 4 * field)) & 15; }

// Offsets in doubles inside one draw's arena; the int tables start at
// ibase (in doubles) and their offsets are in ints from there.
struct ArenaLayout {
    int64_t stride;                  // doubles per draw
    int64_t face, bnd, ring, q, qprev;
    int64_t obs, onb;                // [nobs][kNO] probe records, [nnb][3] stencil dx, dy, w
    int64_t ibase;                   // start of the int tables (doubles)
    int64_t islot, iface, ibflag;    // in ints: [nc][8] slot codes then neighbours,
                                     // [nf][2] owner/neighbour, [nb] boundary flags
    int64_t iobs, ionb;              // in ints: [nobs][8] type|component<<4, cell,
                                     // stencil begin/end, friction faces a/b; [nnb][2] cell, face
    int64_t image_bytes;             // bytes of the pre-filled image copied at launch
};

// Algebraic multigrid hierarchy for the pressure correction: patterns and
// transfer maps shared by every draw on the mesh (built on the host once);
// each draw's values live in its workspace at the given offsets (doubles).
constexpr int kMaxAmgLevels = 8;
constexpr int kMaxAmgCoarse = 128;  // rows of the coarsest level (dense inverse)
struct AmgLevel {
    int n, nnz;
    const int *rowptr, *col, *diagpos;    // CSR pattern (levels >= 1)
    const int *agg;                       // [n] row -> aggregate on the next level
    const int *mptr, *midx;               // next level's rows -> members here, ascending
    const int *gptr, *gidx;               // next level's entries -> entries here (Galerkin)
    int64_t val, x, b, r, dinv;           // workspace offsets (levels >= 1)
};
struct AmgMesh {
    int levels;                           // 0: no AMG
    int coarse_n;                         // rows of the last level
    int64_t inv;                          // workspace offset of its dense inverse
    int64_t workspace;                    // doubles per draw
    const int *src0;                      // level-0 CSR entry -> cell*8 + (0 diag | 1+slot)
    AmgLevel lev[kMaxAmgLevels];
};

// Mesh and problem tables in device memory, shared by every draw on the mesh.
struct DevMesh {
    // pre-filled arena image (geometry, zero state, int tables), twice: cell,
    // face and boundary records as records (img, for arenas in LDS) and as
    // one array per variable (img_soa, for arenas in HBM, where a wave's
    // loads of one variable of 64 consecutive cells then share cache lines)
    const double *img;
    const double *img_soa;
    int nc, nif, nf, nb, npatch;
    int has_outlet, planar;
    const double *wall_y;        // [nc] wall-face distance of the constrained cell, NaN otherwise
    const double *folded;        // [4 nc] (area, delta) x 2 of folded front/back faces, or null
    int nobs;                    // prepared observations (their stencils are in the image)
    ArenaLayout lay;  // per-draw arena layout for this mesh
    AmgMesh amg;      // pressure-correction multigrid (levels == 0 when unused)
};

struct DevSettings {
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
    int sst_variant;
    double nu;
    double init_ux, init_uy, init_p, init_k, init_w;
    int mixed_precision;   // 1: FP32 Krylov with FP64 defect correction
    int profile;           // 1: accumulate per-phase clock counts in DevResult::cycles
    int grid_blocks;       // workgroups per draw (1: one-workgroup draws)
    int amg_block_rows;    // multigrid levels of at most this many rows: one workgroup
    int grid_plain_launch; // 1: multi-workgroup grids by an ordinary dispatch (host side only)
};

// One draw to solve: its mesh, its arena and its coefficients.
struct DevSlot {
    const DevMesh *mesh;
    double *arena;          // used when the arena lives in global memory
    double *amg;            // multigrid workspace (mesh.amg.workspace doubles) or null
    const double *warm;     // flow state to start from (warm start) or null
    double *save;           // where to write the final flow state, or null
    double coeffs[11];
    double tol;             // convergence tolerance of this row (0: the settings')
    int max_iter;           // iteration cap of this row (0: the settings')
    int flags;              // kSlot* bits
    int row;                // output row index
};
// slot flags: no observation convergence monitor (a solver without
// setConvergenceMonitor, as ParameterSensitivity's); re-apply the boundary
// conditions to a warm state before solving (ParameterSensitivity.cpp:790-794)
constexpr int kSlotNoMonitor = 1, kSlotWarmBCs = 2;

// A flow state (warm start source and export), in doubles: the interior
// fields Ux, Uy, p, k, omega, nuT, F1, F2, Pk, CDkw (nc each), the boundary
// values Ux, Uy, p, k, omega (max(nb, 1) each), then the turbulence-
// established flag (FlowFields, core/include/FlowFields.hpp:12-35).
constexpr int kStateInterior = 10, kStateBoundary = 5;
inline int64_t state_doubles(int nc, int nb) {
    return (int64_t)kStateInterior * nc + (int64_t)kStateBoundary * (nb > 1 ? nb : 1) + 2;
}

// Per-draw results written by the kernel.
struct DevResult {
    int status, iters, stalled, converged, diverged, error;
    double worst, rate;
    long long krylov[5];
    long long cycles[kPhases];
};

// ---------------------------------------------------------------- serving
constexpr int kServeBlock = 256;      // threads per serving workgroup
constexpr int kServeBlockLarge = 512; // ... when a registered mesh has more than 256 cells
constexpr int kMaxServeObs = 64;      // observations per served problem

// A registered problem: its mesh tables, settings and arena placement.
struct DevProblem {
    DevMesh mesh;
    DevSettings settings;
    int lds;  // 1: the arena fits the workgroup's LDS
};

// Host-written control block (pinned host memory).
struct ServeControl {
    long long tail;  // number of requests published
    int stop;        // 1: workgroups return once the published requests are claimed
    int pad;
};

// One request of the ring (pinned host memory).
struct ServeRequest {
    long long ticket;
    int problem;
    int pad;
    double coeffs[11];
};

// One result of the ring (pinned host memory); done = ticket + 1 is written
// last, after the result.
struct ServeOutput {
    DevResult res;
    double predictions[kMaxServeObs];
    long long done;
};

}  // namespace rq
