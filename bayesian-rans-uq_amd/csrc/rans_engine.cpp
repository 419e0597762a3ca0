// Host side of the MI355X forward-solve engine: admission of the problem,
// mesh tables for the cell-centric kernels, device memory, launch and the
// per-row result conventions of the reference's ForwardBatch
// (rans-uq/include/ForwardBatch.hpp:52-124) and ForwardModel
// (rans-uq/src/ForwardModel.cpp:35-137).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "../../include/rans_amd.h"
#include "rans_device.h"
#include "warm_cache.h"

namespace rq {
hipError_t launch_simple_solve(int block, bool lds, const DevSlot *slots, int nslots,
                               const DevSettings &s, DevResult *results, double *pred,
                               int pred_stride, double *fields, int64_t field_stride,
                               size_t lds_bytes, hipStream_t stream);
hipError_t launch_serve(int block, int grid, const DevProblem *problems, ServeControl *ctl,
                        const ServeRequest *ring, ServeOutput *outs, int ring_size,
                        unsigned long long *head, double *arenas, int64_t arena_stride,
                        double *amgs, int64_t amg_stride, long long spin_limit, size_t lds_bytes,
                        hipStream_t stream);
int grid_co_resident(size_t lds_bytes, bool mixed);

// meshes from this size up run multi-workgroup draws when the batch leaves CUs idle
constexpr int kGridAutoCells = 4096;
hipError_t launch_grid_solve(const DevSlot *slots, int nslots, int nblk, const DevSettings &s,
                             DevResult *results, double *pred, int pred_stride, double *fields,
                             int64_t field_stride, GridState *states, size_t lds_bytes,
                             hipStream_t stream);
hipError_t noalloc_scratch(void **ptr, size_t *bytes);
std::mutex &noalloc_lock(int dev);
hipError_t launch_fill(double *p, int64_t n, double v, hipStream_t stream);
}

using namespace rq;

namespace {

void set_err(char *err, size_t cap, const std::string &msg) {
    if (err && cap) {
        std::snprintf(err, cap, "%s", msg.c_str());
    }
}

struct Fail {
    int code;
    std::string msg;
};

#define HIPCHECK(expr)                                                                      \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            throw Fail{e_ == hipErrorOutOfMemory ? RQ_ENOMEM : RQ_EDEVICE,                  \
                       std::string(#expr) + ": " + hipGetErrorString(e_)};                  \
    } while (0)

// Device buffer holding many host arrays back to back.
struct Blob {
    std::vector<char> host;
    template <class T>
    size_t add(const std::vector<T> &v) {
        size_t off = (host.size() + 255) & ~size_t(255);
        host.resize(off + v.size() * sizeof(T) + 1);
        if (!v.empty()) std::memcpy(host.data() + off, v.data(), v.size() * sizeof(T));
        return off;
    }
};

bool finite3(const double *v) {
    return std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2]);
}

}  // namespace

struct rq_engine {
    // owned problem copy
    int nc = 0, nif = 0, nf = 0, nb = 0, npatch = 0, K = 0;
    int n_active = 0, nobs = 0;
    bool wall_crossings = false;  // separation or reattachment observations present
    bool drag = false;            // drag observations present
    bool folded = false;          // front/back faces of a one-layer mesh folded into the rows
    std::vector<int> active;
    std::vector<double> base, lower, upper, observed, variance;
    DevSettings settings{};
    int block = 64;
    bool lds = true;
    size_t lds_bytes = 0;
    ArenaLayout lay{};
    DevMesh mesh_host{};  // device pointers, host copy of the table
    // device
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    char *d_blob = nullptr;
    DevMesh *d_mesh = nullptr;
    DevSlot *d_slots = nullptr;
    DevResult *d_results = nullptr;
    double *d_pred = nullptr, *d_fields = nullptr, *d_arena = nullptr, *d_amg = nullptr;
    GridState *d_grid = nullptr;
    size_t cap_grid = 0;
    int cus = 0;
    bool block_user = false;  // block_size given explicitly
    int last_nblk = 1;
    size_t cap_slots = 0, cap_results = 0, cap_pred = 0, cap_fields = 0, cap_arena = 0;
    size_t cap_amg = 0;
    double last_ms = 0.0;
    // hipGraph replay of one-workgroup-per-draw batch launches: the launch
    // (slot upload, solve kernel, result download through pinned staging
    // buffers) is captured once and replayed while its key stays the same
    struct GraphKey {
        int ns = 0, block = 0, nobs = 0;
        bool lds = false, fields = false;
        size_t lds_bytes = 0;
        const void *ptr[8] = {};
        DevSettings set{};  // kernel argument baked into the graph
        bool operator==(const GraphKey &o) const {
            if (ns != o.ns || block != o.block || nobs != o.nobs || lds != o.lds ||
                fields != o.fields || lds_bytes != o.lds_bytes ||
                std::memcmp(&set, &o.set, sizeof(DevSettings)) != 0)
                return false;
            for (int i = 0; i < 8; ++i)
                if (ptr[i] != o.ptr[i]) return false;
            return true;
        }
    };
    bool graph_on = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    GraphKey gkey;
    DevSlot *h_slots = nullptr;
    DevResult *h_results = nullptr;
    double *h_pred = nullptr, *h_fields = nullptr;
    size_t hcap_slots = 0, hcap_results = 0, hcap_pred = 0, hcap_fields = 0;
    int64_t graph_launches = 0;
    // serialises the entry points that use the engine's buffers and stream,
    // so evaluate may be called from several threads (one call at a time)
    mutable std::recursive_mutex mu;

    void drop_graph() {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        gexec = nullptr;
        graph = nullptr;
    }

    ~rq_engine() {
        // nothing of this engine may still run when its memory goes back
        if (stream) (void)hipStreamSynchronize(stream);
        drop_graph();
        if (h_slots) (void)hipHostFree(h_slots);
        if (h_results) (void)hipHostFree(h_results);
        if (h_pred) (void)hipHostFree(h_pred);
        if (h_fields) (void)hipHostFree(h_fields);
        if (d_blob) (void)hipFree(d_blob);
        if (d_mesh) (void)hipFree(d_mesh);
        if (d_slots) (void)hipFree(d_slots);
        if (d_results) (void)hipFree(d_results);
        if (d_pred) (void)hipFree(d_pred);
        if (d_fields) (void)hipFree(d_fields);
        if (d_arena) (void)hipFree(d_arena);
        if (d_amg) (void)hipFree(d_amg);
        if (d_grid) (void)hipFree(d_grid);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

void validate_settings(const rq_settings &s) {
    // SolverSettings::validate (core/include/FlowFields.hpp:203-240)
    if (s.sst_variant < 0 || s.sst_variant > 2)
        throw Fail{RQ_EINVAL, "unknown SST variant: expected Full (0), NoLimiter (1), or KOmega (2)"};
    if (s.max_iterations < 0 || s.inner_iterations <= 0 || s.convergence_window < 2 ||
        s.monitor_interval <= 0 || s.turb_update_interval <= 0 || s.turb_start_iter < 0 ||
        s.nut_floor_iters < 0 ||
        s.turb_start_iter > std::numeric_limits<int>::max() - s.nut_floor_iters)
        throw Fail{RQ_EINVAL, "invalid SIMPLE iteration count or cadence"};
    if (!std::isfinite(s.divergence_limit) || s.divergence_limit < 0.0)
        throw Fail{RQ_EINVAL, "SIMPLE divergence limit must be finite and nonnegative"};
    if (!std::isfinite(s.convergence_tol) || s.convergence_tol < 0.0)
        throw Fail{RQ_EINVAL, "SIMPLE convergence tolerance must be finite and nonnegative"};
    for (double v : {s.inner_tolerance, s.alpha_u, s.alpha_k, s.alpha_omega})
        if (!std::isfinite(v) || v <= 0.0)
            throw Fail{RQ_EINVAL,
                       "SIMPLE tolerances and relaxation denominators must be finite and positive"};
    if (!std::isfinite(s.alpha_p) || s.alpha_p < 0.)
        throw Fail{RQ_EINVAL, "SIMPLE pressure/source relaxation must be finite and nonnegative"};
    if (!std::isfinite(s.body_force_x) || !std::isfinite(s.body_force_y))
        throw Fail{RQ_EINVAL, "SIMPLE body force must be finite"};
    for (double v : {s.k_min, s.k_max, s.omega_min})
        if (!std::isfinite(v) || v < 0.)
            throw Fail{RQ_EINVAL, "SIMPLE turbulence limits must be finite and nonnegative"};
    if (s.k_max < s.k_min) throw Fail{RQ_EINVAL, "SIMPLE kMax must be at least kMin"};
    for (int v : {s.momentum_convection, s.turbulence_convection})
        if (v < 0 || v > 2) throw Fail{RQ_EINVAL, "unknown SIMPLE convection scheme"};
    for (int v : {s.pressure_solver, s.momentum_solver, s.turbulence_solver})
        if (v < 0 || v > 2) throw Fail{RQ_EINVAL, "unknown linear solver"};
    if (s.momentum_solver == SOLVER_AMG || s.turbulence_solver == SOLVER_AMG)
        throw Fail{RQ_EINVAL, "AMG is supported for the pressure correction only"};
    if (s.convergence_window > kMaxWindow)
        throw Fail{RQ_EINVAL, "convergence window exceeds the engine limit of 256"};
    if (s.mixed_precision != 0 && s.mixed_precision != 1)
        throw Fail{RQ_EINVAL, "mixed_precision must be 0 or 1"};
    if (s.profile != 0 && s.profile != 1) throw Fail{RQ_EINVAL, "profile must be 0 or 1"};
    if (s.grid_blocks < 0 || s.grid_blocks > kMaxGridBlocks)
        throw Fail{RQ_EINVAL, "grid_blocks must be in [0, 64]"};
    if (s.amg_block_rows < 0) throw Fail{RQ_EINVAL, "amg_block_rows must be nonnegative"};
    if (s.block_size != 0 && s.block_size != 64 && s.block_size != 128 && s.block_size != 256 &&
        s.block_size != 512)
        throw Fail{RQ_EINVAL, "block_size must be 0, 64, 128, 256 or 512"};
}

// preparePlanarOmegaBalance (rans-uq/src/SIMPLESolver.cpp:92-153)
bool planar_balance(const rq_mesh_desc &m, const std::vector<int> &bpatch, const rq_bc_desc &bc,
                    std::vector<double> &sing, std::vector<double> &lap, std::vector<double> &grad) {
    const int nif = m.n_internal, nf = m.n_faces, nc = m.n_cells;
    if (nf == nif) return false;
    const int f0 = nif;
    const double inx = m.face_normal[3 * f0] * -1.0, iny = m.face_normal[3 * f0 + 1] * -1.0;
    auto dotc = [&](const double *v) { return v[0] * inx + v[1] * iny; };
    const double origin = dotc(&m.face_center[3 * f0]);
    double height = 0;
    bool periodic = false;
    for (int f = nif; f < nf; ++f) {
        const int pi = bpatch[f - nif];
        if (m.patch_type[pi] != PATCH_WALL || bc.w_type[pi] != BC_WALL_KOMEGA ||
            std::abs(std::abs(dotc(&m.face_normal[3 * f])) - 1.0) > 1e-12)
            return false;
        height = std::max(height, dotc(&m.face_center[3 * f]) - origin);
    }
    if (!(height > 0)) return false;
    const double tolerance = 1e-12 * height;
    for (int f = 0; f < nf; ++f) {
        const double *d = &m.face_d[3 * f], *n = &m.face_normal[3 * f], *s = &m.face_shift[3 * f];
        const double de = m.face_delta[f];
        const double ex = d[0] - n[0] * de, ey = d[1] - n[1] * de, ez = d[2] - n[2] * de;
        if (std::sqrt(ex * ex + ey * ey + ez * ez) > 1e-9 * de || std::abs(dotc(s)) > tolerance)
            return false;
        periodic = periodic || (s[0] * s[0] + s[1] * s[1] + s[2] * s[2]) > 0;
        if (f >= nif) {
            const double y = dotc(&m.face_center[3 * f]) - origin;
            if (std::min(std::abs(y), std::abs(y - height)) > tolerance) return false;
            if ((y < height / 2 ? -1.0 : 1.0) * dotc(n) < 1 - 1e-12) return false;
        }
    }
    if (!periodic) return false;
    sing.assign(nc, 0.0);
    lap.assign(nc, 0.0);
    grad.assign(std::max(nif, 1), 0.0);
    for (int c = 0; c < nc; ++c) {
        const double y = dotc(&m.cell_center[3 * c]) - origin;
        if (!(y > 0 && y < height)) return false;
        sing[c] = 1 / (y * y) + 1 / ((height - y) * (height - y));
        lap[c] = 6 * (std::pow(y, -4) + std::pow(height - y, -4));
    }
    for (int f = 0; f < nif; ++f) {
        const double y = dotc(&m.face_center[3 * f]) - origin;
        if (!(y > 0 && y < height)) return false;
        grad[f] = 2 * (std::pow(height - y, -3) - std::pow(y, -3)) * dotc(&m.face_normal[3 * f]);
    }
    return true;
}

int64_t build_layout(ArenaLayout &L, int nc, int nif, int nf, int nb, int nobs, int nnb,
                     int window) {
    (void)nif;
    int64_t o = (int64_t)nc * kNA;
    L.face = o;
    o += (int64_t)nf * kNF;
    L.bnd = o;
    o += (int64_t)std::max(nb, 1) * kNB;
    L.ring = o;
    o += 8 * (int64_t)window;
    L.q = o;
    o += std::max(nobs, 1);
    L.qprev = o;
    o += std::max(nobs, 1);
    L.obs = o;
    o += (int64_t)std::max(nobs, 1) * kNO;
    L.onb = o;
    o += 3 * (int64_t)std::max(nnb, 1);
    o = (o + 1) & ~int64_t(1);
    L.ibase = o;
    int64_t ni = 0;
    L.islot = ni;
    ni += 8 * (int64_t)nc;
    L.iface = ni;
    ni += 2 * (int64_t)nf;
    L.ibflag = ni;
    ni += std::max(nb, 1);
    L.iobs = ni;
    ni += 8 * (int64_t)std::max(nobs, 1);
    L.ionb = ni;
    ni += 2 * (int64_t)std::max(nnb, 1);
    o += (ni + 1) / 2;
    o = (o + 1) & ~int64_t(1);  // 16-byte multiple
    L.stride = o;
    L.image_bytes = o * 8;
    return o;
}

// ---------------------------------------------------------------------------
// Multigrid hierarchy of the pressure correction, built once per mesh.
// Level 0 is the LDU system in CSR form with each row sorted by column
// (lduToCSR, core/include/LinearSolver.hpp:472-532); its values for the
// setup are the geometric Laplacian A_f / delta_f (the pressure-correction
// coefficients without the per-draw (V/aP)_f factor, plus the outlet faces
// and the cell-0 pin). Aggregates follow buildAggregates
// (LinearSolver.hpp:535-575): greedy seeds with strength
// |a_ij| / sqrt(|a_ii a_jj|) > 0.25, orphans joining their strongest
// neighbour. Coarse patterns follow galerkinProduct (LinearSolver.hpp:578-618)
// and coarsening stops at 50 rows or when it stalls (LinearSolver.hpp:797-812).
// Only the finest level uses the strength filter: on coarser levels every
// unaggregated neighbour joins the seed, because the filtered coarsening
// stalls on anisotropic near-wall meshes (a 120x64 hill keeps 470 of 7680
// rows after 15 levels) while a V-cycle in one workgroup needs few levels
// and a small dense coarsest solve. Each level then aggregates its
// aggregates once more (no filter): the filtered pass mostly pairs rows
// along the strong direction, and every level of a multi-workgroup V-cycle
// costs grid barriers. Measured: 120x64 hill 7680/3845/1701/447/121/34 rows
// -> 7680/1701/121/34, 32 -> 35 CG iterations per p' solve, 4.4 -> 3.5 ms per
// outer iteration; 480x256 eight levels -> five, 27.7 -> 20.8 ms. From 16384
// cells up the finest level aggregates a third time (480x256: 20.8 -> 19.8 ms,
// 240x128: 7.3 -> 6.3 ms; on the 120x64 hill it costs 3.5 -> 3.6 ms, as the
// CG iterations grow from 35 to 42).
// ---------------------------------------------------------------------------
struct AmgHostLevel {
    int n = 0;
    std::vector<int> rowptr, col, diagpos, agg, mptr, midx, gptr, gidx;
    std::vector<double> val;
};

struct AmgHost {
    std::vector<AmgHostLevel> lev;
    std::vector<int> src0;
};

// meshes from this many cells up aggregate their finest level three times
constexpr int kAmgThirdPassRows = 16384;

void amg_aggregate(const AmgHostLevel &A, double threshold, std::vector<int> &map,
                   int &ncoarse) {
    const int n = A.n;
    map.assign(n, -1);
    ncoarse = 0;
    std::vector<double> diag(n, 0.0);
    for (int i = 0; i < n; ++i) diag[i] = A.val[A.diagpos[i]];
    for (int i = 0; i < n; ++i) {
        if (map[i] >= 0) continue;
        const int agg = ncoarse++;
        map[i] = agg;
        const double aii = std::abs(diag[i]);
        for (int j = A.rowptr[i]; j < A.rowptr[i + 1]; ++j) {
            const int c = A.col[j];
            if (c == i || map[c] >= 0) continue;
            const double ajj = std::abs(diag[c]);
            const double strength = std::abs(A.val[j]) / std::sqrt(aii * ajj + 1e-30);
            if (strength > threshold) map[c] = agg;
        }
    }
    for (int i = 0; i < n; ++i) {
        if (map[i] >= 0) continue;
        double best = -1.0;
        int bestAgg = 0;
        for (int j = A.rowptr[i]; j < A.rowptr[i + 1]; ++j) {
            const int c = A.col[j];
            if (c == i && map[c] < 0) continue;
            if (map[c] >= 0 && std::abs(A.val[j]) > best) {
                best = std::abs(A.val[j]);
                bestAgg = map[c];
            }
        }
        map[i] = bestAgg;
    }
}

void amg_build(const rq_mesh_desc &m, const int *ISL, const int *IBF, bool has_outlet,
               AmgHost &H) {
    const int nc = m.n_cells, nif = m.n_internal;
    AmgHostLevel L0;
    L0.n = nc;
    H.src0.clear();
    L0.rowptr.assign(nc + 1, 0);
    for (int c = 0; c < nc; ++c) {
        struct E {
            int col, src;
            double v;
        };
        std::vector<E> row;
        double d = 0.0;
        for (int s = 0; s < kSlots; ++s) {
            const int code = ISL[8 * c + s];
            if (code < 0) continue;
            const int f = code & 0x3fffffff;
            const double g = m.face_area[f] / std::max(m.face_delta[f], 1e-20);
            if (f < nif) {
                const int nbr = ISL[8 * c + 4 + s];
                if (nbr == c) throw Fail{RQ_EINVAL, "the AMG pressure solver needs a mesh "
                                                    "without self-periodic faces"};
                row.push_back({nbr, 1 + s, -g});
                d += g;
            } else if (IBF[f - nif] & kFlagOutlet) {
                d += g;
            }
        }
        if (!has_outlet && c == 0) d += std::abs(d) > 1e-30 ? d : 1.0;
        row.push_back({c, 0, d});
        std::sort(row.begin(), row.end(), [](const E &a, const E &b) { return a.col < b.col; });
        for (size_t k = 1; k < row.size(); ++k)
            if (row[k].col == row[k - 1].col)
                throw Fail{RQ_EINVAL, "the AMG pressure solver needs rows without repeated "
                                      "columns"};
        for (const E &x : row) {
            if (x.col == c) L0.diagpos.push_back((int)L0.col.size());
            L0.col.push_back(x.col);
            L0.val.push_back(x.v);
            H.src0.push_back(c * 8 + x.src);
        }
        L0.rowptr[c + 1] = (int)L0.col.size();
    }
    H.lev.clear();
    H.lev.push_back(std::move(L0));
    for (int l = 0; l + 1 < kMaxAmgLevels; ++l) {
        AmgHostLevel &F = H.lev[l];
        if (F.n <= 50) break;
        std::vector<int> map;
        int ncoarse = 0;
        amg_aggregate(F, l == 0 ? 0.25 : 0.0, map, ncoarse);
        if (ncoarse == 0 || ncoarse >= F.n) break;
        // large meshes merge their finest aggregates a third time
        const int passes = (l == 0 && F.n >= kAmgThirdPassRows) ? 3 : 2;
        for (int pass = 1; pass < passes && ncoarse > 50; ++pass) {
            // a second pass over the aggregates (their Galerkin graph, no
            // strength filter) merges them again: coarse rows gather about
            // four fine rows instead of two, so the hierarchy has fewer levels
            AmgHostLevel T;
            T.n = ncoarse;
            std::vector<std::map<int, double>> trows(ncoarse);
            for (int i = 0; i < F.n; ++i)
                for (int j = F.rowptr[i]; j < F.rowptr[i + 1]; ++j)
                    trows[map[i]][map[F.col[j]]] += F.val[j];
            T.rowptr.assign(ncoarse + 1, 0);
            T.diagpos.assign(ncoarse, -1);
            for (int I = 0; I < ncoarse; ++I) {
                for (const auto &kv : trows[I]) {
                    if (kv.first == I) T.diagpos[I] = (int)T.col.size();
                    T.col.push_back(kv.first);
                    T.val.push_back(kv.second);
                }
                T.rowptr[I + 1] = (int)T.col.size();
            }
            bool diag_ok = true;
            for (int I = 0; I < ncoarse; ++I) diag_ok = diag_ok && T.diagpos[I] >= 0;
            std::vector<int> map2;
            int n2 = 0;
            if (diag_ok) amg_aggregate(T, 0.0, map2, n2);
            if (!(n2 > 0 && n2 < ncoarse)) break;
            for (int &a : map) a = map2[a];
            ncoarse = n2;
        }
        // Galerkin pattern, each coarse row sorted by column
        std::vector<std::map<int, double>> rows(ncoarse);
        for (int i = 0; i < F.n; ++i)
            for (int j = F.rowptr[i]; j < F.rowptr[i + 1]; ++j) rows[map[i]][map[F.col[j]]] += F.val[j];
        // Coarse rows are padded to a multiple of 4 entries with (column =
        // row, value 0) entries after the sorted real ones, so the device
        // reads a row 4 entries at a time; the padding never receives a
        // Galerkin contribution and adds exact zeros.
        AmgHostLevel Cl;
        Cl.n = ncoarse;
        Cl.rowptr.assign(ncoarse + 1, 0);
        Cl.diagpos.assign(ncoarse, -1);
        std::vector<int> nreal(ncoarse, 0);
        for (int I = 0; I < ncoarse; ++I) {
            for (const auto &kv : rows[I]) {
                if (kv.first == I) Cl.diagpos[I] = (int)Cl.col.size();
                Cl.col.push_back(kv.first);
                Cl.val.push_back(kv.second);
            }
            nreal[I] = (int)Cl.col.size() - Cl.rowptr[I];
            while ((Cl.col.size() - Cl.rowptr[I]) % 4) {
                Cl.col.push_back(I);
                Cl.val.push_back(0.0);
            }
            Cl.rowptr[I + 1] = (int)Cl.col.size();
        }
        for (int I = 0; I < ncoarse; ++I)
            if (Cl.diagpos[I] < 0) throw Fail{RQ_EINVAL, "AMG coarse level without a diagonal"};
        // fine entry -> coarse entry, gathered per coarse entry in fine order
        std::vector<std::vector<int>> contrib(Cl.col.size());
        for (int i = 0; i < F.n; ++i)
            for (int j = F.rowptr[i]; j < F.rowptr[i + 1]; ++j) {
                const int I = map[i], J = map[F.col[j]];
                const int *b = Cl.col.data() + Cl.rowptr[I], *e2 = b + nreal[I];
                const int pos = (int)(std::lower_bound(b, e2, J) - Cl.col.data());
                contrib[pos].push_back(j);
            }
        F.gptr.assign(Cl.col.size() + 1, 0);
        for (size_t k = 0; k < contrib.size(); ++k) {
            F.gidx.insert(F.gidx.end(), contrib[k].begin(), contrib[k].end());
            F.gptr[k + 1] = (int)F.gidx.size();
        }
        F.agg = map;
        std::vector<std::vector<int>> members(ncoarse);
        for (int i = 0; i < F.n; ++i) members[map[i]].push_back(i);
        // member lists padded to a multiple of 4 with -1 (read 4 at a time)
        F.mptr.assign(ncoarse + 1, 0);
        for (int I = 0; I < ncoarse; ++I) {
            F.midx.insert(F.midx.end(), members[I].begin(), members[I].end());
            while ((F.midx.size() - F.mptr[I]) % 4) F.midx.push_back(-1);
            F.mptr[I + 1] = (int)F.midx.size();
        }
        H.lev.push_back(std::move(Cl));
    }
    if (H.lev.back().n > kMaxAmgCoarse)
        throw Fail{RQ_EINVAL, "AMG coarsening stalled above 128 rows"};
}

rq_engine *create_impl(const rq_mesh_desc *mp, const rq_param_desc *pp, const rq_obs_desc *op,
                       const rq_bc_desc *bp, double nu, const rq_settings *sp, const double *init) {
    if (!mp || !pp || !op || !bp || !sp || !init) throw Fail{RQ_EINVAL, "null argument"};
    const rq_mesh_desc &m = *mp;
    const rq_settings &s = *sp;
    validate_settings(s);
    // simple_admission::problem / initial (rans-uq/include/SIMPLEValidation.hpp:32-60)
    if (!std::isfinite(nu) || nu <= 0)
        throw Fail{RQ_EINVAL, "SIMPLE kinematic viscosity must be finite and positive"};
    if (!std::isfinite(init[0]) || !std::isfinite(init[1]) || !std::isfinite(init[2]))
        throw Fail{RQ_EINVAL, "SIMPLE initial velocity and pressure must be finite"};
    if (!std::isfinite(init[3]) || init[3] < 0)
        throw Fail{RQ_EINVAL, "SIMPLE initial k must be finite and nonnegative"};
    if (!std::isfinite(init[4]) || init[4] <= 0)
        throw Fail{RQ_EINVAL, "SIMPLE initial omega must be finite and positive"};
    const int nc = m.n_cells, nif = m.n_internal, nf = m.n_faces, np = m.n_patches;
    if (nc <= 0 || nif < 0 || nf < nif || np < 0) throw Fail{RQ_EINVAL, "invalid mesh counts"};
    const int nb = nf - nif;
    for (int f = 0; f < nf; ++f) {
        const bool internal = f < nif;
        if (m.owner[f] < 0 || m.owner[f] >= nc ||
            (internal ? (m.neighbor[f] < 0 || m.neighbor[f] >= nc) : m.neighbor[f] != -1))
            throw Fail{RQ_EINVAL, "invalid mesh face connectivity"};
        if (m.face_normal[3 * f + 2] != 0.0)
            throw Fail{RQ_EINVAL, "mesh must be 2D: every face normal needs a zero z component"};
        if (!finite3(&m.face_normal[3 * f]) || !finite3(&m.face_center[3 * f]) ||
            !finite3(&m.face_d[3 * f]) || !finite3(&m.face_shift[3 * f]) ||
            !std::isfinite(m.face_area[f]) || !std::isfinite(m.face_delta[f]) ||
            !std::isfinite(m.face_weight[f]))
            throw Fail{RQ_EINVAL, "nonfinite mesh face geometry"};
    }
    // each cell's faces ascending and every face listed by its cells
    std::vector<int> seen(nf, 0);
    int K = 0;
    for (int c = 0; c < nc; ++c) {
        const int a = m.cell_face_ptr[c], b = m.cell_face_ptr[c + 1];
        if (b < a) throw Fail{RQ_EINVAL, "invalid cell face list"};
        K = std::max(K, b - a);
        for (int q = a; q < b; ++q) {
            const int f = m.cell_face_idx[q];
            if (f < 0 || f >= nf) throw Fail{RQ_EINVAL, "invalid cell face index"};
            if (q > a && f < m.cell_face_idx[q - 1])
                throw Fail{RQ_EINVAL, "cell face lists must be in ascending face order"};
            if (m.owner[f] != c && m.neighbor[f] != c)
                throw Fail{RQ_EINVAL, "inconsistent cell-face incidence"};
            ++seen[f];
        }
        if (!std::isfinite(m.cell_volume[c]) || !finite3(&m.cell_center[3 * c]) ||
            !std::isfinite(m.wall_distance[c]))
            throw Fail{RQ_EINVAL, "nonfinite mesh cell geometry or wall distance"};
    }
    for (int f = 0; f < nf; ++f)
        if (seen[f] != (f < nif ? 2 : 1)) throw Fail{RQ_EINVAL, "incomplete cell face lists"};
    if (K > kSlots) throw Fail{RQ_EINVAL, "cells with more than 4 faces are not supported"};
    if (m.cell_folded)
        for (int64_t i = 0; i < 2 * (int64_t)nc; ++i) {
            const double a = m.cell_folded[2 * i], d = m.cell_folded[2 * i + 1];
            if (!std::isfinite(a) || a < 0 || (a > 0 && !(std::isfinite(d) && d > 0)))
                throw Fail{RQ_EINVAL, "invalid folded front/back face geometry"};
        }
    // boundary faces: patches must list them in face order
    std::vector<int> bpatch(nb, -1);
    {
        int expect = nif;
        for (int pi = 0; pi < np; ++pi)
            for (int q = m.patch_face_ptr[pi]; q < m.patch_face_ptr[pi + 1]; ++q) {
                if (m.patch_face_idx[q] != expect)
                    throw Fail{RQ_EINVAL,
                               "patches must list the boundary faces in ascending face order"};
                bpatch[expect - nif] = pi;
                ++expect;
            }
        if (expect != nf) throw Fail{RQ_EINVAL, "every boundary face needs a patch"};
    }
    // boundary conditions (validatePatchBCs, BoundaryCondition.hpp:45-103)
    const rq_bc_desc &bc = *bp;
    for (int pi = 0; pi < np; ++pi) {
        for (const int *t : {bc.u_type, bc.p_type, bc.k_type, bc.w_type})
            if (t[pi] < 0 || t[pi] > 8) throw Fail{RQ_EINVAL, "unknown boundary condition type"};
        const int ut = bc.u_type[pi];
        if ((ut == BC_DIRICHLET || ut == BC_INLET || ut == BC_WALL_MOVING) &&
            (!std::isfinite(bc.u_val[2 * pi]) || !std::isfinite(bc.u_val[2 * pi + 1])))
            throw Fail{RQ_EINVAL, "active vector boundary values must be finite"};
        if ((bc.p_type[pi] == BC_DIRICHLET || bc.p_type[pi] == BC_OUTLET) &&
            !std::isfinite(bc.p_val[pi]))
            throw Fail{RQ_EINVAL, "active scalar boundary values must be finite"};
        for (auto pr : {std::make_pair(bc.k_type, bc.k_val), std::make_pair(bc.w_type, bc.w_val)})
            if (pr.first[pi] == BC_DIRICHLET) {
                if (!std::isfinite(pr.second[pi]))
                    throw Fail{RQ_EINVAL, "active scalar boundary values must be finite"};
                if (pr.second[pi] < 0)
                    throw Fail{RQ_EINVAL, "SIMPLE prescribed turbulence must be nonnegative"};
            }
    }
    // parameters (InferenceParameterSet::validate)
    const rq_param_desc &P = *pp;
    if (P.n_active <= 0)
        throw Fail{RQ_EINVAL, "inference parameters require at least one active index"};
    {
        bool used[11] = {};
        for (int i = 0; i < P.n_active; ++i) {
            const int a = P.active_indices[i];
            if (a < 0 || a >= 11) throw Fail{RQ_EINVAL, "inference parameter index must be in [0,10]"};
            if (used[a]) throw Fail{RQ_EINVAL, "inference parameter indices must be unique"};
            used[a] = true;
        }
    }

    auto *E = new rq_engine();
    try {
        E->nc = nc;
        E->nif = nif;
        E->nf = nf;
        E->nb = nb;
        E->npatch = np;
        E->K = K;
        E->n_active = P.n_active;
        E->active.assign(P.active_indices, P.active_indices + P.n_active);
        E->base.assign(P.base, P.base + 11);
        E->lower.assign(P.lower, P.lower + P.n_active);
        E->upper.assign(P.upper, P.upper + P.n_active);
        const rq_obs_desc &O = *op;
        E->nobs = O.n_obs;
        E->observed.assign(O.observed, O.observed + O.n_obs);
        E->variance.assign(O.variance, O.variance + O.n_obs);
        for (int i = 0; i < O.n_obs; ++i)
            if (!std::isfinite(E->variance[i]) || E->variance[i] <= 0 ||
                !std::isfinite(E->observed[i]))
                throw Fail{RQ_EINVAL,
                           "Gaussian measurements require finite values, nonnegative uncertainties "
                           "and finite positive total variance"};

        // ---- arena image: geometry in the cell/face records and the int
        //      tables, each face constant computed with the reference's arithmetic
        const int no = O.n_obs;
        const int nnb = no > 0 ? O.nb_ptr[no] : 0;
        for (int i = 0; i < no; ++i) {
            const int t = O.type[i];
            if (t < 0 || t > 5) throw Fail{RQ_EINVAL, "unsupported observation type"};
            const bool scan = t == OBS_DRAG || t == OBS_SEPARATION || t == OBS_REATTACHMENT;
            if (t == OBS_DRAG) E->drag = true;
            if (t == OBS_SEPARATION || t == OBS_REATTACHMENT) E->wall_crossings = true;
            if (scan)
                for (int q = O.nb_ptr[i]; q < O.nb_ptr[i + 1]; ++q)
                    if (O.nb_cell[q] >= 0)
                        throw Fail{RQ_EINVAL, "wall-crossing stencil must list boundary faces"};
            if ((t == OBS_VELOCITY || t == OBS_PRESSURE) && (O.cell[i] < 0 || O.cell[i] >= nc))
                throw Fail{RQ_EINVAL, "observation probe cell out of range"};
            if (O.component[i] < 0 || O.component[i] > 2)
                throw Fail{RQ_EINVAL, "observation component out of range"};
            if (O.nb_ptr[i + 1] < O.nb_ptr[i] || O.nb_ptr[i] < 0)
                throw Fail{RQ_EINVAL, "invalid observation stencil"};
            for (int q = O.nb_ptr[i]; q < O.nb_ptr[i + 1]; ++q)
                if (O.nb_cell[q] >= nc || (O.nb_cell[q] < 0 && (O.nb_face[q] < nif || O.nb_face[q] >= nf)))
                    throw Fail{RQ_EINVAL, "observation stencil index out of range"};
            if (t == OBS_SKIN_FRICTION && (O.face_a[i] >= nf || O.face_b[i] >= nf))
                throw Fail{RQ_EINVAL, "skin-friction stencil face out of range"};
        }
        ArenaLayout L{};
        build_layout(L, nc, nif, nf, nb, no, nnb, s.convergence_window);
        std::vector<double> image(L.stride, 0.0);
        double *CR = image.data(), *FR = image.data() + L.face;
        int *IB = reinterpret_cast<int *>(image.data() + L.ibase);
        int *ISL = IB + L.islot, *IFA = IB + L.iface, *IBF = IB + L.ibflag;
        for (int64_t i = 0; i < 8 * (int64_t)nc; ++i) ISL[i] = -1;
        for (int c = 0; c < nc; ++c) {
            int s2 = 0, lastSelf = -1;
            for (int q = m.cell_face_ptr[c]; q < m.cell_face_ptr[c + 1]; ++q, ++s2) {
                const int f = m.cell_face_idx[q];
                int side = 0;
                if (f < nif) {
                    if (m.owner[f] == m.neighbor[f]) {  // self-wrapped face: owner side first
                        side = (lastSelf == f) ? 1 : 0;
                        lastSelf = f;
                    } else {
                        side = m.owner[f] == c ? 0 : 1;
                    }
                    ISL[8 * c + 4 + s2] = side == 0 ? m.neighbor[f] : m.owner[f];
                }
                ISL[8 * c + s2] = f | (side << 30);
            }
        }
        for (int f = 0; f < nf; ++f) {
            IFA[2 * f] = m.owner[f];
            IFA[2 * f + 1] = m.neighbor[f];
        }
        for (int b = 0; b < nb; ++b) {
            const int pi = bpatch[b];
            const int t = m.patch_type[pi];
            IBF[b] = (t == PATCH_WALL ? kFlagWall : 0) | (t == PATCH_OUTLET ? kFlagOutlet : 0) |
                     (bc.u_type[pi] << 4) | (bc.p_type[pi] << 8) | (bc.k_type[pi] << 12) |
                     (bc.w_type[pi] << 16);
            double *R = image.data() + L.bnd + (int64_t)b * kNB;
            R[B_uvx] = bc.u_val[2 * pi];
            R[B_uvy] = bc.u_val[2 * pi + 1];
            R[B_pv] = bc.p_val[pi];
            R[B_kv] = bc.k_val[pi];
            R[B_wv] = bc.w_val[pi];
        }
        // prepared observations: one record per probe plus its stencil entries
        for (int i = 0; i < no; ++i) {
            double *R = image.data() + L.obs + (int64_t)i * kNO;
            R[O_offx] = O.offx[i];
            R[O_offy] = O.offy[i];
            R[O_axx] = O.axx[i];
            R[O_axy] = O.axy[i];
            R[O_ayy] = O.ayy[i];
            R[O_det] = O.det[i];
            R[O_t] = O.t[i];
            R[O_refv] = O.ref_velocity[i];
            R[O_refp] = O.ref_pressure[i];
            R[O_refrho] = O.ref_density[i];
            R[O_floor] = std::sqrt(O.variance[i]);
            int *Q = IB + L.iobs + 8 * (int64_t)i;
            Q[0] = O.type[i] | (O.component[i] << 4);
            Q[1] = O.cell[i];
            Q[2] = O.nb_ptr[i];
            Q[3] = O.nb_ptr[i + 1];
            Q[4] = O.face_a[i];
            Q[5] = O.face_b[i];
        }
        for (int q = 0; q < nnb; ++q) {
            double *R = image.data() + L.onb + 3 * (int64_t)q;
            R[0] = O.nb_dx[q];
            R[1] = O.nb_dy[q];
            R[2] = O.nb_w[q];
            IB[L.ionb + 2 * q] = O.nb_cell[q];
            IB[L.ionb + 2 * q + 1] = O.nb_face[q];
        }
        const double *CC = m.cell_center, *FC = m.face_center, *FN = m.face_normal;
        const double *FSH = m.face_shift, *FD = m.face_d;
        const bool linear = s.momentum_convection == CONV_LINEAR_UPWIND ||
                            s.turbulence_convection == CONV_LINEAR_UPWIND;
        const bool limited = s.momentum_convection == CONV_LIMITED ||
                             s.turbulence_convection == CONV_LIMITED;
        if (linear && limited)
            throw Fail{RQ_EINVAL, "momentum and turbulence must use the same higher-order scheme"};
        for (int f = 0; f < nf; ++f) {
            double *R = FR + (int64_t)f * kNF;
            const double area = m.face_area[f];
            R[W_Sx] = FN[3 * f] * area;  // Vec3 Sf = normal * area
            R[W_Sy] = FN[3 * f + 1] * area;
            R[W_A] = area;
            R[W_dlt] = std::max(m.face_delta[f], 1e-20);
            R[W_w] = m.face_weight[f];
            R[W_wm] = 1.0 - m.face_weight[f];
            R[W_nx] = FN[3 * f];
            R[W_ny] = FN[3 * f + 1];
            const double dd = std::max(m.face_delta[f], 1e-30);
            R[W_ex] = FD[3 * f] / dd;
            R[W_ey] = FD[3 * f + 1] / dd;
            if (f < nif) {
                const int o = m.owner[f], n = m.neighbor[f];
                if (linear) {
                    // x_f - x_U (SIMPLESolver.cpp:517-518)
                    R[W_dOx] = FC[3 * f] - CC[3 * o];
                    R[W_dOy] = FC[3 * f + 1] - CC[3 * o + 1];
                    R[W_dNx] = (FC[3 * f] - FSH[3 * f]) - CC[3 * n];
                    R[W_dNy] = (FC[3 * f + 1] - FSH[3 * f + 1]) - CC[3 * n + 1];
                } else {
                    // x_D - x_U of the limited scheme (SIMPLESolver.cpp:521-523)
                    R[W_dOx] = (CC[3 * n] + FSH[3 * f]) - CC[3 * o];
                    R[W_dOy] = (CC[3 * n + 1] + FSH[3 * f + 1]) - CC[3 * o + 1];
                    R[W_dNx] = (CC[3 * o] - FSH[3 * f]) - CC[3 * n];
                    R[W_dNy] = (CC[3 * o + 1] - FSH[3 * f + 1]) - CC[3 * n + 1];
                }
            }
        }
        std::vector<double> sing, lap, grad;
        const bool planar = planar_balance(m, bpatch, bc, sing, lap, grad);
        for (int f = 0; f < nif && planar; ++f) FR[(int64_t)f * kNF + W_omg] = grad[f];
        for (int c = 0; c < nc; ++c) {
            double *R = CR + (int64_t)c * kNA;
            R[V_vol] = m.cell_volume[c];
            R[V_oms] = planar ? sing[c] : 0.0;
            R[V_oml] = planar ? lap[c] : 0.0;
            R[V_wd] = m.wall_distance[c];
        }
        // wall-constrained omega cells: the last wall face of each owner wins
        // (SIMPLESolver::omegaWallValues, SIMPLESolver.cpp:665-684)
        std::vector<double> wall_y(nc, std::numeric_limits<double>::quiet_NaN());
        for (int pi = 0; pi < np; ++pi) {
            if (m.patch_type[pi] != PATCH_WALL) continue;
            for (int q = m.patch_face_ptr[pi]; q < m.patch_face_ptr[pi + 1]; ++q) {
                const int f = m.patch_face_idx[q];
                wall_y[m.owner[f]] = std::max(m.face_delta[f], 1e-20);
            }
        }
        bool has_outlet = false;
        for (int pi = 0; pi < np; ++pi)
            if (m.patch_type[pi] == PATCH_OUTLET && m.patch_face_ptr[pi + 1] > m.patch_face_ptr[pi])
                has_outlet = true;

        AmgHost H;
        const bool use_amg = s.pressure_solver == SOLVER_AMG;
        if (use_amg) amg_build(m, ISL, IBF, has_outlet, H);

        // ---- device copy: the arena image, the initial wall-omega table and
        //      the multigrid patterns
        Blob B;
        DevMesh D{};
        const size_t o_img = B.add(image);
        // the same image with the cell, face and boundary records stored as
        // one array per variable, for arenas in HBM
        std::vector<double> soa(image);
        const int nbs = std::max(nb, 1);
        for (int c = 0; c < nc; ++c)
            for (int v = 0; v < kNA; ++v)
                soa[(int64_t)v * nc + c] = image[(int64_t)c * kNA + v];
        for (int f = 0; f < nf; ++f)
            for (int v = 0; v < kNF; ++v)
                soa[L.face + (int64_t)v * nf + f] = image[L.face + (int64_t)f * kNF + v];
        for (int b = 0; b < nbs; ++b)
            for (int v = 0; v < kNB; ++v)
                soa[L.bnd + (int64_t)v * nbs + b] = image[L.bnd + (int64_t)b * kNB + v];
        const size_t o_soa = B.add(soa);
        size_t o_wy = B.add(wall_y);
        size_t o_fold = 0;
        if (m.cell_folded)
            o_fold = B.add(std::vector<double>(m.cell_folded, m.cell_folded + 4 * (int64_t)nc));
        struct LevOff {
            size_t rowptr, col, diagpos, agg, mptr, midx, gptr, gidx;
        };
        std::vector<LevOff> lo(H.lev.size());
        size_t o_src0 = 0;
        if (use_amg) {
            o_src0 = B.add(H.src0);
            for (size_t l = 0; l < H.lev.size(); ++l) {
                const AmgHostLevel &A = H.lev[l];
                lo[l] = {B.add(A.rowptr), B.add(A.col), B.add(A.diagpos), B.add(A.agg),
                         B.add(A.mptr),   B.add(A.midx), B.add(A.gptr),  B.add(A.gidx)};
            }
        }

        HIPCHECK(hipGetDevice(&E->device));
        HIPCHECK(hipMalloc(&E->d_blob, B.host.size()));
        HIPCHECK(hipMemcpy(E->d_blob, B.host.data(), B.host.size(), hipMemcpyHostToDevice));
        char *b0 = E->d_blob;
        auto Dd = [&](size_t o) { return reinterpret_cast<const double *>(b0 + o); };
        D.img = Dd(o_img);
        D.img_soa = Dd(o_soa);
        D.nc = nc; D.nif = nif; D.nf = nf; D.nb = nb; D.npatch = np;
        D.has_outlet = has_outlet; D.planar = planar;
        D.wall_y = Dd(o_wy);
        D.folded = m.cell_folded ? Dd(o_fold) : nullptr;
        E->folded = m.cell_folded != nullptr;
        D.nobs = no;
        if (use_amg) {
            auto I = [&](size_t o) { return reinterpret_cast<const int *>(b0 + o); };
            AmgMesh &G = D.amg;
            G.levels = (int)H.lev.size();
            G.src0 = I(o_src0);
            int64_t w = 0;
            for (int l = 0; l < G.levels; ++l) {
                const AmgHostLevel &A = H.lev[l];
                AmgLevel &g = G.lev[l];
                g.n = A.n;
                g.nnz = (int)A.col.size();
                g.rowptr = I(lo[l].rowptr);
                g.col = I(lo[l].col);
                g.diagpos = I(lo[l].diagpos);
                g.agg = I(lo[l].agg);
                g.mptr = I(lo[l].mptr);
                g.midx = I(lo[l].midx);
                g.gptr = I(lo[l].gptr);
                g.gidx = I(lo[l].gidx);
                if (l == 0) {
                    g.val = g.x = g.b = g.r = g.dinv = -1;
                    continue;
                }
                g.val = w;
                w += g.nnz;
                g.x = w;
                w += g.n;
                g.b = w;
                w += g.n;
                g.r = w;
                w += g.n;
                g.dinv = w;
                w += g.n;
            }
            G.coarse_n = H.lev.back().n;
            G.inv = w;
            w += 2 * (int64_t)G.coarse_n * G.coarse_n + G.coarse_n;
            G.workspace = (w + 1) & ~int64_t(1);
        }
        D.lay = L;
        E->lay = L;
        E->mesh_host = D;
        HIPCHECK(hipMalloc(&E->d_mesh, sizeof(DevMesh)));
        HIPCHECK(hipMemcpy(E->d_mesh, &D, sizeof(DevMesh), hipMemcpyHostToDevice));

        // ---- settings and launch shape
        DevSettings &S = E->settings;
        S.max_iterations = s.max_iterations;
        S.convergence_tol = s.convergence_tol;
        S.divergence_limit = s.divergence_limit;
        S.convergence_window = s.convergence_window;
        S.monitor_interval = s.monitor_interval;
        S.alpha_u = s.alpha_u;
        S.alpha_p = s.alpha_p;
        S.alpha_k = s.alpha_k;
        S.alpha_omega = s.alpha_omega;
        S.inner_iterations = s.inner_iterations;
        S.inner_tolerance = s.inner_tolerance;
        S.turb_start_iter = s.turb_start_iter;
        S.turb_update_interval = s.turb_update_interval;
        S.nut_floor_iters = s.nut_floor_iters;
        S.momentum_convection = s.momentum_convection;
        S.turbulence_convection = s.turbulence_convection;
        S.body_force_x = s.body_force_x;
        S.body_force_y = s.body_force_y;
        S.k_min = s.k_min;
        S.k_max = s.k_max;
        S.omega_min = s.omega_min;
        S.pressure_solver = s.pressure_solver;
        S.momentum_solver = s.momentum_solver;
        S.turbulence_solver = s.turbulence_solver;
        S.sst_variant = s.sst_variant;
        S.nu = nu;
        S.init_ux = init[0];
        S.init_uy = init[1];
        S.init_p = init[2];
        S.init_k = init[3];
        S.init_w = init[4];
        S.mixed_precision = s.mixed_precision;
        S.profile = s.profile;
        S.grid_blocks = s.grid_blocks;
        S.amg_block_rows = s.amg_block_rows > 0 ? s.amg_block_rows : 1024;
        S.grid_plain_launch = s.grid_plain_launch != 0;
        HIPCHECK(hipDeviceGetAttribute(&E->cus, hipDeviceAttributeMultiprocessorCount, E->device));
        E->block_user = s.block_size != 0;
        if (s.block_size) E->block = s.block_size;
        else E->block = nc <= 64 ? 64 : nc <= 256 ? 256 : 512;
        int maxLds = 0;
        HIPCHECK(hipDeviceGetAttribute(&maxLds, hipDeviceAttributeMaxSharedMemoryPerBlock,
                                       E->device));
        const size_t arenaBytes = (size_t)E->lay.stride * sizeof(double);
        const size_t staticLds = 4096;
        E->lds = arenaBytes + staticLds <= (size_t)maxLds;
        E->lds_bytes = E->lds ? arenaBytes : 0;
        HIPCHECK(hipStreamCreateWithFlags(&E->stream, hipStreamNonBlocking));
        HIPCHECK(hipEventCreate(&E->ev0));
        HIPCHECK(hipEventCreate(&E->ev1));
    } catch (...) {
        delete E;
        throw;
    }
    return E;
}

template <class T>
void grow(T *&ptr, size_t &cap, size_t n) {
    if (n <= cap) return;
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    cap = 0;
    HIPCHECK(hipMalloc(&ptr, n * sizeof(T)));
    cap = n;
}

// pinned host staging buffer of at least n elements (a new buffer has a new
// address, which invalidates a captured graph through its key)
template <class T>
void grow_pinned(T *&ptr, size_t &cap, size_t n) {
    if (n <= cap) return;
    if (ptr) HIPCHECK(hipHostFree(ptr));
    ptr = nullptr;
    cap = 0;
    if (hipHostMalloc((void **)&ptr, n * sizeof(T), hipHostMallocDefault) != hipSuccess) {
        ptr = nullptr;
        throw Fail{RQ_ENOMEM, "pinned host allocation failed"};
    }
    cap = n;
}

void write_row_error(const rq_batch_out *out, int64_t row, const char *msg) {
    if (out->row_errors && out->row_error_stride)
        std::snprintf(out->row_errors + row * out->row_error_stride, out->row_error_stride, "%s",
                      msg);
}

// Status, log-likelihood and error text of one solved row
// (ForwardModel.cpp:92-135, ObservationOperator.hpp:592-607).
// msg is a string literal or null (no heap text, so the allocation-free
// evaluation can use it too)
struct RowFinal {
    int status;
    double loglik;
    bool keep_predictions;
    const char *msg;
};

RowFinal finalize_row(const rq_engine *E, const DevResult &d, const double *pred) {
    RowFinal f{d.status, -1e6, false, nullptr};
    if (f.status == ST_UNKNOWN)
        f.msg = d.error == ERR_NONFINITE_MONITOR ? "convergence monitor returned a nonfinite observation"
                : d.error == ERR_GRID_TIMEOUT   ? "device synchronisation of the draw timed out"
                                                : "linear system coefficients and source must be finite";
    if (f.status == ST_CONVERGED || f.status == ST_UNCONVERGED) {
        double acc = 0.0;
        for (int j = 0; j < E->nobs; ++j) {
            const double pj = pred[j];
            if (!std::isfinite(pj)) {
                f.msg = "Likelihood predictions must be finite";
                break;
            }
            const double diff = pj - E->observed[j];
            acc += -0.5 * diff * diff / E->variance[j];
            if (!std::isfinite(acc)) {
                f.msg = "Gaussian log-likelihood exceeds finite range";
                break;
            }
        }
        if (!f.msg) {
            f.loglik = acc;
            f.keep_predictions = true;
        } else {
            f.status = ST_UNKNOWN;
        }
    }
    return f;
}

// InferenceParameterSet::inBounds and SSTCoefficients::validate for one row;
// fills the 11 coefficients. Returns 0, or ST_INVALID / ST_UNKNOWN with text.
int admit_row(const rq_engine *E, const double *row, double *coeffs, const char *&msg) {
    for (int i = 0; i < E->n_active; ++i) {
        const double v = row[i];
        if (v < E->lower[i] || v > E->upper[i] || std::isnan(v) || std::isinf(v)) {
            msg = "coefficient row is nonfinite or out of bounds";
            return ST_INVALID;
        }
    }
    for (int i = 0; i < 11; ++i) coeffs[i] = E->base[i];
    for (int i = 0; i < E->n_active; ++i) coeffs[E->active[i]] = row[i];
    bool ok = true;
    for (int i = 0; i < 11; ++i)
        if (!std::isfinite(coeffs[i]) || coeffs[i] < 0) ok = false;
    if (!ok || coeffs[2] == 0. || coeffs[8] == 0.) {
        msg = "SST coefficients must be finite and nonnegative, with positive beta1 and betaStar";
        return ST_UNKNOWN;
    }
    return 0;
}

// Dynamic LDS of a multi-workgroup launch: each workgroup's bookkeeping
// replica (four contraction rings of 2m samples, the current and the previous
// monitor values).
size_t grid_lds_bytes(const rq_engine *E) {
    return (8 * (size_t)E->settings.convergence_window + 2 * (size_t)std::max(E->nobs, 1)) * 8;
}

// Workgroups per draw for a launch of ns admitted rows.
// Automatic selection (grid_blocks 0, no explicit block size): arenas in
// HBM of at least kGridAutoCells cells spread each draw over the CUs the
// batch leaves idle, one workgroup per 256 cells at most. Below that size
// one workgroup per draw is faster (measured: 60x32 hill 2.3 ms per outer
// iteration on one workgroup against 2.8 on eight; 120x64 9.3 against
// 4.7 on sixteen).
int choose_grid_blocks(const rq_engine *E, int ns, size_t grid_lds) {
    int nblk = 1;
    if (E->settings.grid_blocks > 1) {
        nblk = E->settings.grid_blocks;
    } else if (E->settings.grid_blocks == 0 && !E->lds && !E->block_user &&
               E->nc >= kGridAutoCells) {
        const int per_xcd = (ns + kXcds - 1) / kXcds;
        nblk = std::min({kMaxGridBlocks, E->cus / (kXcds * per_xcd), (E->nc + 255) / 256});
    }
    if (nblk > 1) {
        // every workgroup of the padded grid (draws spread over the XCDs) must
        // be co-resident
        const int co = grid_co_resident(grid_lds, E->settings.mixed_precision != 0);
        const int per_xcd = (ns + kXcds - 1) / kXcds;
        const int fit = co / (kXcds * per_xcd);
        nblk = std::min(nblk, fit);
        if (nblk < 2) nblk = 1;
    }
    // the multi-workgroup kernels do not evaluate drag or fold front/back
    // faces, and the mixed-precision one evaluates neither separation nor
    // reattachment: such problems run one workgroup per draw
    if (nblk > 1 && (E->drag || E->folded ||
                     (E->settings.mixed_precision != 0 && E->wall_crossings)))
        nblk = 1;
    return nblk;
}

// Per-row solve options of the sensitivity path: tolerance and iteration
// cap (0: the engine's settings) and kSlot* flags.
struct RowOpts {
    double tol = 0.0;
    int max_iter = 0;
    int flags = 0;
};

// warm/save: per input row, the device flow state a row starts from and the
// device buffer its final state is written to (either may be null); opts:
// per input row solve options, or null
int evaluate_impl(rq_engine *E, const double *theta, int64_t n_rows, int32_t n_params,
                  int32_t workers, rq_batch_out *out, const double *const *warm = nullptr,
                  double *const *save = nullptr, const RowOpts *opts = nullptr) {
    if (!E || !out || !out->status) throw Fail{RQ_EINVAL, "null argument"};
    std::lock_guard<std::recursive_mutex> lock(E->mu);
    if (workers < 1) throw Fail{RQ_EINVAL, "workers must be positive"};
    if (n_params != E->n_active)
        throw Fail{RQ_EINVAL, "batch rows must all have n_parameters entries"};
    if (n_rows < 0) throw Fail{RQ_EINVAL, "row count must be nonnegative"};
    if (n_rows > 0 && !theta) throw Fail{RQ_EINVAL, "null coefficient rows"};
    const double nan = std::numeric_limits<double>::quiet_NaN();
    const int nobs = E->nobs;
    std::vector<int64_t> valid;
    std::vector<DevSlot> slots;
    for (int64_t r = 0; r < n_rows; ++r) {
        const double *row = theta + r * n_params;
        if (out->row_errors && out->row_error_stride) out->row_errors[r * out->row_error_stride] = 0;
        if (out->predictions)
            for (int i = 0; i < nobs; ++i) out->predictions[r * nobs + i] = nan;
        if (out->iterations) out->iterations[r] = 0;
        if (out->final_worst_measure) out->final_worst_measure[r] = 0.0;
        if (out->contraction_rate_u) out->contraction_rate_u[r] = 0.0;
        if (out->stalled) out->stalled[r] = 0;
        if (out->converged) out->converged[r] = 0;
        if (out->krylov_iterations)
            for (int e = 0; e < 5; ++e) out->krylov_iterations[r * 5 + e] = 0;
        if (out->phase_cycles)
            for (int e = 0; e < kPhases; ++e) out->phase_cycles[r * kPhases + e] = 0;
        // InferenceParameterSet::inBounds (core/include/InferenceParameters.hpp:127-140)
        // and SSTCoefficients::validate (core/include/SSTModel.hpp:29-45)
        DevSlot sl{};
        sl.mesh = E->d_mesh;
        const char *msg = nullptr;
        const int bad = admit_row(E, row, sl.coeffs, msg);
        if (bad) {
            out->status[r] = bad;
            if (out->loglik) out->loglik[r] = bad == ST_INVALID ? -1e30 : -1e6;
            write_row_error(out, r, msg);
            continue;
        }
        sl.row = (int)slots.size();
        sl.warm = warm ? warm[r] : nullptr;
        sl.save = save ? save[r] : nullptr;
        if (opts) {
            sl.tol = opts[r].tol;
            sl.max_iter = opts[r].max_iter;
            sl.flags = opts[r].flags;
        }
        valid.push_back(r);
        slots.push_back(sl);
    }
    if (slots.empty()) return RQ_OK;
    const int ns = (int)slots.size();
    HIPCHECK(hipSetDevice(E->device));
    grow(E->d_slots, E->cap_slots, ns);
    grow(E->d_results, E->cap_results, ns);
    grow(E->d_pred, E->cap_pred, (size_t)ns * std::max(nobs, 1));
    const int64_t fstride = 10 * (int64_t)E->nc + E->nf;
    if (out->fields) grow(E->d_fields, E->cap_fields, (size_t)ns * fstride);
    if (!E->lds) {
        grow(E->d_arena, E->cap_arena, (size_t)ns * E->lay.stride);
        for (int i = 0; i < ns; ++i) slots[i].arena = E->d_arena + (size_t)i * E->lay.stride;
    }
    const int64_t amgw = E->mesh_host.amg.workspace;
    if (amgw > 0) {
        grow(E->d_amg, E->cap_amg, (size_t)ns * amgw);
        for (int i = 0; i < ns; ++i) slots[i].amg = E->d_amg + (size_t)i * amgw;
    }
    // workgroups per draw: meshes whose arena lives in HBM spread each draw
    // over several co-resident workgroups when the batch leaves CUs idle
    const size_t grid_lds = grid_lds_bytes(E);
    int nblk = choose_grid_blocks(E, ns, grid_lds);
    if (nblk > 1 && E->lds) {
        // the grid mode keeps the arena in HBM
        grow(E->d_arena, E->cap_arena, (size_t)ns * E->lay.stride);
        for (int i = 0; i < ns; ++i) slots[i].arena = E->d_arena + (size_t)i * E->lay.stride;
    }
    E->last_nblk = nblk;
    const size_t npred = (size_t)ns * std::max(nobs, 1);
    std::vector<DevResult> res;
    std::vector<double> pred, fields;
    const DevResult *resp = nullptr;
    const double *predp = nullptr, *fieldsp = nullptr;
    if (E->graph_on && nblk == 1) {
        // graph replay: this call's slots go through the pinned staging buffer
        // the captured upload reads; results come back through pinned buffers
        grow_pinned(E->h_slots, E->hcap_slots, ns);
        grow_pinned(E->h_results, E->hcap_results, ns);
        grow_pinned(E->h_pred, E->hcap_pred, npred);
        if (out->fields) grow_pinned(E->h_fields, E->hcap_fields, (size_t)ns * fstride);
        std::memcpy(E->h_slots, slots.data(), ns * sizeof(DevSlot));
        rq_engine::GraphKey k;
        k.ns = ns;
        k.block = E->block;
        k.nobs = nobs;
        k.lds = E->lds;
        k.fields = out->fields != nullptr;
        k.lds_bytes = E->lds_bytes;
        std::memcpy(&k.set, &E->settings, sizeof(DevSettings));
        const void *ptrs[8] = {E->d_slots, E->d_results, E->d_pred, E->d_fields,
                               E->h_slots,  E->h_results, E->h_pred, E->h_fields};
        for (int i = 0; i < 8; ++i) k.ptr[i] = ptrs[i];
        if (!E->gexec || !(k == E->gkey)) {
            E->drop_graph();
            HIPCHECK(hipStreamBeginCapture(E->stream, hipStreamCaptureModeThreadLocal));
            hipError_t e = hipMemcpyAsync(E->d_slots, E->h_slots, ns * sizeof(DevSlot),
                                          hipMemcpyHostToDevice, E->stream);
            if (e == hipSuccess)
                e = launch_simple_solve(E->block, E->lds, E->d_slots, ns, E->settings,
                                        E->d_results, E->d_pred, std::max(nobs, 1),
                                        out->fields ? E->d_fields : nullptr, fstride,
                                        E->lds_bytes, E->stream);
            if (e == hipSuccess)
                e = hipMemcpyAsync(E->h_results, E->d_results, ns * sizeof(DevResult),
                                   hipMemcpyDeviceToHost, E->stream);
            if (e == hipSuccess)
                e = hipMemcpyAsync(E->h_pred, E->d_pred, npred * sizeof(double),
                                   hipMemcpyDeviceToHost, E->stream);
            if (e == hipSuccess && out->fields)
                e = hipMemcpyAsync(E->h_fields, E->d_fields, (size_t)ns * fstride * sizeof(double),
                                   hipMemcpyDeviceToHost, E->stream);
            // the capture always ends, also after a failed step
            hipGraph_t g = nullptr;
            const hipError_t e2 = hipStreamEndCapture(E->stream, &g);
            E->graph = g;
            if (e == hipSuccess) e = e2;
            if (e == hipSuccess) e = hipGraphInstantiate(&E->gexec, E->graph, nullptr, nullptr, 0);
            if (e != hipSuccess) {
                E->drop_graph();
                throw Fail{RQ_EDEVICE, std::string("graph capture of the batch launch: ") +
                                           hipGetErrorString(e)};
            }
            E->gkey = k;
        }
        HIPCHECK(hipEventRecord(E->ev0, E->stream));
        HIPCHECK(hipGraphLaunch(E->gexec, E->stream));
        HIPCHECK(hipEventRecord(E->ev1, E->stream));
        HIPCHECK(hipStreamSynchronize(E->stream));
        E->graph_launches++;
        resp = E->h_results;
        predp = E->h_pred;
        fieldsp = E->h_fields;
    } else {
        HIPCHECK(hipMemcpyAsync(E->d_slots, slots.data(), ns * sizeof(DevSlot),
                                hipMemcpyHostToDevice, E->stream));
        if (nblk > 1) {
            grow(E->d_grid, E->cap_grid, ns);
            HIPCHECK(hipMemsetAsync(E->d_grid, 0, ns * sizeof(GridState), E->stream));
        }
        HIPCHECK(hipEventRecord(E->ev0, E->stream));
        if (nblk > 1)
            HIPCHECK(launch_grid_solve(E->d_slots, ns, nblk, E->settings, E->d_results,
                                       E->d_pred, std::max(nobs, 1),
                                       out->fields ? E->d_fields : nullptr, fstride, E->d_grid,
                                       grid_lds, E->stream));
        else
            HIPCHECK(launch_simple_solve(E->block, E->lds, E->d_slots, ns, E->settings,
                                         E->d_results, E->d_pred, std::max(nobs, 1),
                                         out->fields ? E->d_fields : nullptr, fstride,
                                         E->lds_bytes, E->stream));
        HIPCHECK(hipEventRecord(E->ev1, E->stream));
        res.resize(ns);
        pred.resize(npred);
        HIPCHECK(hipMemcpyAsync(res.data(), E->d_results, ns * sizeof(DevResult),
                                hipMemcpyDeviceToHost, E->stream));
        HIPCHECK(hipMemcpyAsync(pred.data(), E->d_pred, pred.size() * sizeof(double),
                                hipMemcpyDeviceToHost, E->stream));
        if (out->fields) {
            fields.resize((size_t)ns * fstride);
            HIPCHECK(hipMemcpyAsync(fields.data(), E->d_fields, fields.size() * sizeof(double),
                                    hipMemcpyDeviceToHost, E->stream));
        }
        HIPCHECK(hipStreamSynchronize(E->stream));
        resp = res.data();
        predp = pred.data();
        fieldsp = fields.data();
    }
    float ms = 0.f;
    HIPCHECK(hipEventElapsedTime(&ms, E->ev0, E->ev1));
    E->last_ms = ms;
    for (int i = 0; i < ns; ++i) {
        const int64_t r = valid[i];
        const DevResult &d = resp[i];
        const RowFinal fin = finalize_row(E, d, predp + (size_t)i * std::max(nobs, 1));
        const int status = fin.status;
        const char *msg = fin.msg;
        if (fin.keep_predictions && out->predictions)
            for (int j = 0; j < nobs; ++j)
                out->predictions[r * nobs + j] = predp[(size_t)i * nobs + j];
        out->status[r] = status;
        if (out->loglik) out->loglik[r] = fin.loglik;
        if (out->iterations) out->iterations[r] = d.iters;
        if (out->final_worst_measure) out->final_worst_measure[r] = d.worst;
        if (out->contraction_rate_u) out->contraction_rate_u[r] = d.rate;
        if (out->stalled) out->stalled[r] = (uint8_t)d.stalled;
        if (out->converged) out->converged[r] = status == ST_CONVERGED;
        if (out->krylov_iterations)
            for (int e = 0; e < 5; ++e) out->krylov_iterations[r * 5 + e] = d.krylov[e];
        if (out->phase_cycles)
            for (int e = 0; e < kPhases; ++e) out->phase_cycles[r * kPhases + e] = d.cycles[e];
        if (msg) write_row_error(out, r, msg);
        if (out->fields)
            std::memcpy(out->fields + r * fstride, fieldsp + (size_t)i * fstride,
                        fstride * sizeof(double));
    }
    return RQ_OK;
}

}  // namespace

// ----------------------------------------------------------------------------
// Single-row model with a warm-start cache: ForwardModel
// (rans-uq/include/ForwardModel.hpp:191-247, rans-uq/src/ForwardModel.cpp:35-137).
// Converged flow states stay in device memory; a warm-started row copies the
// nearest cached state into its arena instead of the uniform initialisation.
// ----------------------------------------------------------------------------
struct rq_model {
    rq_engine *E = nullptr;
    rq::WarmCache cache;
    int64_t state_n = 0;              // doubles per flow state
    std::vector<double *> pool;       // device state buffers not in the cache
    std::vector<double *> owned;      // every device state buffer
    explicit rq_model(const rq::WarmCache &c) : cache(c) {}
    ~rq_model() {
        for (double *p : owned) (void)hipFree(p);
        delete E;
    }
    double *take() {
        if (!pool.empty()) {
            double *p = pool.back();
            pool.pop_back();
            return p;
        }
        double *p = nullptr;
        if (hipMalloc(&p, (size_t)state_n * sizeof(double)) != hipSuccess)
            throw Fail{RQ_ENOMEM, "device allocation of a flow state failed"};
        owned.push_back(p);
        return p;
    }
    void give_back(const std::vector<void *> &freed) {
        for (void *p : freed) pool.push_back(static_cast<double *>(p));
    }
    size_t entry_bytes() const {
        return rq::WarmCache::entry_bytes((size_t)E->n_active, E->nc, E->nb);
    }
};

namespace {

// Rows whose coefficients fail the host checks take no cache lookup
// (ForwardModel.cpp:45-49 returns before it).
bool row_admissible(const rq_engine *E, const double *row) {
    double coeffs[11];
    const char *msg = nullptr;
    return admit_row(E, row, coeffs, msg) == 0;
}

// batch = false: rows one after another, each looking up the cache as it
// stands after the rows before it (the reference's sequence of evaluate
// calls); batch = true: every row looks up the cache as it stands at the
// call, all rows run in one launch, and converged rows are stored in row
// order afterwards.
int model_evaluate_impl(rq_model *M, const double *theta, int64_t n_rows, int32_t n_params,
                        rq_batch_out *out, uint8_t *used, bool batch) {
    if (!M || !out || !out->status) throw Fail{RQ_EINVAL, "null argument"};
    rq_engine *E = M->E;
    if (n_params != E->n_active)
        throw Fail{RQ_EINVAL, "batch rows must all have n_parameters entries"};
    if (n_rows < 0) throw Fail{RQ_EINVAL, "row count must be nonnegative"};
    if (n_rows > 0 && !theta) throw Fail{RQ_EINVAL, "null coefficient rows"};
    HIPCHECK(hipSetDevice(E->device));
    auto after = [&](int64_t r, double *save) {
        std::vector<void *> freed;
        if (out->status[r] == ST_CONVERGED &&
            M->cache.store(theta + r * n_params, (size_t)n_params, save, M->entry_bytes(), freed)) {
        } else {
            freed.push_back(save);
        }
        M->give_back(freed);
    };
    if (!batch) {
        for (int64_t r = 0; r < n_rows; ++r) {
            const double *row = theta + r * n_params;
            const rq::WarmCache::Entry *e =
                row_admissible(E, row) ? M->cache.nearest(row, (size_t)n_params) : nullptr;
            if (used) used[r] = e != nullptr;
            const double *warm = e ? static_cast<const double *>(e->payload) : nullptr;
            double *save = M->take();
            rq_batch_out o1 = *out;
            const int64_t nobs = E->nobs;
            o1.status = out->status + r;
            if (out->loglik) o1.loglik = out->loglik + r;
            if (out->predictions) o1.predictions = out->predictions + r * nobs;
            if (out->iterations) o1.iterations = out->iterations + r;
            if (out->final_worst_measure) o1.final_worst_measure = out->final_worst_measure + r;
            if (out->contraction_rate_u) o1.contraction_rate_u = out->contraction_rate_u + r;
            if (out->stalled) o1.stalled = out->stalled + r;
            if (out->converged) o1.converged = out->converged + r;
            if (out->row_errors && out->row_error_stride)
                o1.row_errors = out->row_errors + r * out->row_error_stride;
            if (out->fields) o1.fields = out->fields + r * (10 * (int64_t)E->nc + E->nf);
            if (out->krylov_iterations) o1.krylov_iterations = out->krylov_iterations + r * 5;
            if (out->phase_cycles) o1.phase_cycles = out->phase_cycles + r * kPhases;
            try {
                evaluate_impl(E, row, 1, n_params, 1, &o1, &warm, &save);
            } catch (...) {
                M->pool.push_back(save);
                throw;
            }
            after(r, save);
        }
        return RQ_OK;
    }
    std::vector<const double *> warm(n_rows, nullptr);
    std::vector<double *> save(n_rows, nullptr);
    for (int64_t r = 0; r < n_rows; ++r) {
        const double *row = theta + r * n_params;
        const rq::WarmCache::Entry *e =
            row_admissible(E, row) ? M->cache.nearest(row, (size_t)n_params) : nullptr;
        if (used) used[r] = e != nullptr;
        warm[r] = e ? static_cast<const double *>(e->payload) : nullptr;
    }
    try {
        for (int64_t r = 0; r < n_rows; ++r) save[r] = M->take();
        evaluate_impl(E, theta, n_rows, n_params, 1, out, warm.data(), save.data());
    } catch (...) {
        for (double *p : save)
            if (p) M->pool.push_back(p);
        throw;
    }
    for (int64_t r = 0; r < n_rows; ++r) after(r, save[r]);
    return RQ_OK;
}

}  // namespace

// ----------------------------------------------------------------------------
// Total sensitivities by warm-started central differences:
// ParameterSensitivity::solveState and ::etaJacobianWarmFD
// (rans-uq/src/ParameterSensitivity.cpp:35-59,729-840). The 22 perturbed
// solves of one Jacobian start from the converged base state, which stays in
// device memory, and run in one launch.
// ----------------------------------------------------------------------------
struct rq_sensitivity {
    rq_engine *E = nullptr;     // all eleven coefficients active
    double *state = nullptr;    // converged base state (device)
    double *scratch = nullptr;  // the next base solve's state (device)
    std::vector<double> theta;  // coefficients of the base state
    bool has_state = false;
    ~rq_sensitivity() {
        if (state) (void)hipFree(state);
        if (scratch) (void)hipFree(scratch);
        delete E;
    }
};

namespace {

// InferenceParameterSet::all11 bounds (core/include/InferenceParameters.hpp:103-122)
const double kAll11Lower[11] = {0.1, 0.1, 0.01, 0.1, 0.1, 0.1, 0.01, 0.1, 0.01, 0.1, 0.35};
const double kAll11Upper[11] = {2.0, 2.0, 0.2, 1.5, 2.0, 2.0, 0.2, 1.5, 0.2, 0.8, 0.50};

rq_batch_out one_row_out(int32_t *st, double *ll, double *pred, int32_t *it, double *worst,
                         double *rate) {
    rq_batch_out o{};
    o.status = st;
    o.loglik = ll;
    o.predictions = pred;
    o.iterations = it;
    o.final_worst_measure = worst;
    o.contraction_rate_u = rate;
    return o;
}

int solve_state_impl(rq_sensitivity *S, const double *theta11, int32_t *status) {
    if (!S || !theta11 || !status) throw Fail{RQ_EINVAL, "null argument"};
    rq_engine *E = S->E;
    S->has_state = false;
    S->theta.clear();
    for (int j = 0; j < 11; ++j)
        if (!(theta11[j] >= kAll11Lower[j] && theta11[j] <= kAll11Upper[j])) {
            *status = ST_INVALID;
            return RQ_OK;
        }
    HIPCHECK(hipSetDevice(E->device));
    const int64_t sn = state_doubles(E->nc, E->nb);
    if (!S->scratch) HIPCHECK(hipMalloc(&S->scratch, (size_t)sn * sizeof(double)));
    const int nobs = std::max(E->nobs, 1);
    std::vector<double> pred(nobs);
    int32_t st = 0, it = 0;
    double ll = 0, worst = 0, rate = 0;
    rq_batch_out o = one_row_out(&st, &ll, pred.data(), &it, &worst, &rate);
    RowOpts opt;
    opt.flags = kSlotNoMonitor;
    double *save = S->scratch;
    evaluate_impl(E, theta11, 1, 11, 1, &o, nullptr, &save, &opt);
    // a solve whose observation fails becomes Unknown here, where the
    // reference's solveState returns Converged (it does not observe)
    *status = st;
    if (st == ST_CONVERGED) {
        std::swap(S->state, S->scratch);
        S->theta.assign(theta11, theta11 + 11);
        S->has_state = true;
    }
    return RQ_OK;
}

int warm_fd_impl(rq_sensitivity *S, const double *theta11, double hRel, double hFloor,
                 int32_t warmMaxIter, double warmTol, const double *absSteps,
                 rq_warm_fd_out *out) {
    if (!S || !theta11 || !out) throw Fail{RQ_EINVAL, "null argument"};
    rq_engine *E = S->E;
    // requireMatchingState (ParameterSensitivity.cpp:61-72)
    if (!S->has_state) throw Fail{RQ_EINVAL, "Sensitivity requires a converged base solveState"};
    if (!std::equal(S->theta.begin(), S->theta.end(), theta11))
        throw Fail{RQ_EINVAL, "Total sensitivity requires the coefficients of the converged base state"};
    if (!std::isfinite(hRel) || hRel <= 0 || !std::isfinite(hFloor) || hFloor <= 0 ||
        warmMaxIter < 0 || !std::isfinite(warmTol) || warmTol < 0)
        throw Fail{RQ_EINVAL, "Invalid warm finite-difference step/tolerance/iteration cap"};
    if (absSteps)
        for (int j = 0; j < 11; ++j)
            if (!std::isfinite(absSteps[j]) || absSteps[j] <= 0)
                throw Fail{RQ_EINVAL, "finite-difference steps require eleven positive finite entries"};
    const int nobs = E->nobs;
    const double nan = std::numeric_limits<double>::quiet_NaN();
    const double inf = std::numeric_limits<double>::infinity();
    for (int i = 0; i < nobs * 11; ++i)
        if (out->d_obs_d_theta) out->d_obs_d_theta[i] = nan;
    for (int j = 0; j < 11; ++j) {
        if (out->log_lik_gradient) out->log_lik_gradient[j] = nan;
        if (out->iterations) out->iterations[j] = 0;
        if (out->worst_measure) out->worst_measure[j] = 0.0;
        if (out->converged) out->converged[j] = 0;
        if (out->positive_status) out->positive_status[j] = ST_UNKNOWN;
        if (out->negative_status) out->negative_status[j] = ST_UNKNOWN;
    }
    if (out->n_residual_evals) *out->n_residual_evals = 0;
    // steps and tolerances (ParameterSensitivity.cpp:769-778,813-822); the
    // rows are + then - for each coefficient
    const double baseTol = E->settings.convergence_tol;
    std::vector<double> rows(22 * 11), h(11), tol(11);
    for (int j = 0; j < 11; ++j) {
        double hj;
        if (absSteps) {
            hj = absSteps[j];
        } else {
            const double hr = (j == 8 || j == 9) ? std::min(hRel, 1e-5) : hRel;
            hj = std::max(hr * std::abs(theta11[j]), hFloor);
        }
        double *tp = &rows[(2 * j) * 11], *tm = &rows[(2 * j + 1) * 11];
        std::copy(theta11, theta11 + 11, tp);
        std::copy(theta11, theta11 + 11, tm);
        tp[j] += hj;
        tm[j] -= hj;
        hj = 0.5 * (tp[j] - tm[j]);
        if (!(hj > 0) || !std::isfinite(hj))
            throw Fail{RQ_EINVAL, "finite-difference step is not representable"};
        h[j] = hj;
        const double rel = hj / std::max(std::abs(theta11[j]), hFloor);
        tol[j] = warmTol > 0 ? warmTol : std::min(baseTol, 0.01 * rel);
        if (out->steps) out->steps[j] = hj;
        if (out->tolerances) out->tolerances[j] = tol[j];
    }
    // rows with a nonpositive or nonfinite coefficient are not solved
    // (ParameterSensitivity.cpp:781-788)
    std::vector<int64_t> solved;
    std::vector<double> thetaS;
    std::vector<RowOpts> opts;
    std::vector<const double *> warm;
    std::vector<int32_t> st(22, ST_INVALID), it(22, 0);
    std::vector<double> worst(22, inf), ll(22, nan), pred(22 * (size_t)std::max(nobs, 1), nan);
    for (int r = 0; r < 22; ++r) {
        const double *t = &rows[r * 11];
        bool ok = true;
        for (int j = 0; j < 11; ++j) ok = ok && std::isfinite(t[j]) && t[j] > 0;
        if (!ok) continue;
        solved.push_back(r);
        thetaS.insert(thetaS.end(), t, t + 11);
        RowOpts o;
        o.tol = tol[r / 2];
        o.max_iter = warmMaxIter;
        o.flags = kSlotNoMonitor | kSlotWarmBCs;
        opts.push_back(o);
        warm.push_back(S->state);
    }
    const int64_t ns = (int64_t)solved.size();
    if (ns > 0) {
        std::vector<int32_t> st2(ns), it2(ns);
        std::vector<double> ll2(ns), w2(ns), rate2(ns), p2(ns * (size_t)std::max(nobs, 1));
        rq_batch_out o = one_row_out(st2.data(), ll2.data(), p2.data(), it2.data(), w2.data(),
                                     rate2.data());
        evaluate_impl(E, thetaS.data(), ns, 11, 1, &o, warm.data(), nullptr, opts.data());
        int64_t evals = 0;
        for (int64_t k = 0; k < ns; ++k) {
            const int64_t r = solved[k];
            st[r] = st2[k];
            it[r] = it2[k];
            evals += it2[k];
            worst[r] = (st2[k] != ST_UNKNOWN && std::isfinite(w2[k])) ? w2[k] : inf;
            ll[r] = ll2[k];
            for (int i = 0; i < nobs; ++i) pred[r * nobs + i] = p2[k * nobs + i];
        }
        if (out->n_residual_evals) *out->n_residual_evals = evals;
    }
    for (int j = 0; j < 11; ++j) {
        const int p = 2 * j, m = 2 * j + 1;
        if (out->positive_status) out->positive_status[j] = st[p];
        if (out->negative_status) out->negative_status[j] = st[m];
        if (out->iterations) out->iterations[j] = it[p] + it[m];
        if (out->worst_measure) out->worst_measure[j] = std::max(worst[p], worst[m]);
        const bool ok = st[p] == ST_CONVERGED && st[m] == ST_CONVERGED;
        if (out->converged) out->converged[j] = ok;
        if (!ok) continue;
        for (int i = 0; i < nobs; ++i)
            if (out->d_obs_d_theta)
                out->d_obs_d_theta[i * 11 + j] = (pred[p * nobs + i] - pred[m * nobs + i]) / (2.0 * h[j]);
        if (out->log_lik_gradient) out->log_lik_gradient[j] = (ll[p] - ll[m]) / (2.0 * h[j]);
    }
    return RQ_OK;
}

template <class F>
int guarded(char *err, size_t errcap, F &&f) {
    try {
        int rc = f();
        set_err(err, errcap, "");
        return rc;
    } catch (const Fail &e) {
        set_err(err, errcap, e.msg);
        return e.code;
    } catch (const std::bad_alloc &) {
        set_err(err, errcap, "host allocation failed");
        return RQ_ENOMEM;
    } catch (const std::exception &e) {
        set_err(err, errcap, e.what());
        return RQ_EDEVICE;
    }
}

}  // namespace

// ----------------------------------------------------------------------------
// Launch-shape autotuner: every candidate workgroup size (and, for arenas in
// HBM, workgroups per draw) runs the caller's rows for a capped number of
// outer iterations; the fastest kernel time is kept for later evaluations.
// The shape changes only the order of parallel reductions.
// ----------------------------------------------------------------------------
namespace {

// The kept shape is the first candidate, in the order tried (fewer threads
// and fewer workgroups first), whose time is within 3 % of the fastest: the
// box-to-box spread of these timings is about 1-2 %, so shapes within noise
// of each other resolve the same way on every device.
constexpr double kTuneSlack = 1.03;

// index of the first finite time within kTuneSlack of the smallest, or -1
int first_within_slack(const double *ms, int n) {
    double fastest = std::numeric_limits<double>::infinity();
    for (int i = 0; i < n; ++i)
        if (std::isfinite(ms[i])) fastest = std::min(fastest, ms[i]);
    if (!std::isfinite(fastest)) return -1;
    for (int i = 0; i < n; ++i)
        if (std::isfinite(ms[i]) && ms[i] <= kTuneSlack * fastest) return i;
    return -1;
}

int autotune_impl(rq_engine *E, const double *theta, int64_t n_rows, int32_t n_params,
                  int32_t iterations, rq_tune_out *out) {
    if (!E || !out || (n_rows > 0 && !theta)) throw Fail{RQ_EINVAL, "null argument"};
    std::lock_guard<std::recursive_mutex> lock(E->mu);
    if (n_rows <= 0) throw Fail{RQ_EINVAL, "the autotuner needs at least one coefficient row"};
    if (n_params != E->n_active)
        throw Fail{RQ_EINVAL, "batch rows must all have n_parameters entries"};
    if (iterations <= 0) throw Fail{RQ_EINVAL, "the autotuner needs a positive iteration cap"};
    struct Cand {
        int block, grid;
    };
    std::vector<Cand> cands;
    for (int b : {64, 128, 256, 512}) cands.push_back({b, 1});
    if (!E->lds && E->nc >= 1024)
        for (int g : {2, 4, 8, 16, 32, 64})
            if (g * 128 <= E->nc) cands.push_back({kGridBlock, g});
    const int saved_block = E->block, saved_grid = E->settings.grid_blocks;
    const bool saved_user = E->block_user;
    std::vector<RowOpts> opts(n_rows);
    for (auto &o : opts) o.max_iter = iterations;
    std::vector<int32_t> st(n_rows), it(n_rows);
    std::vector<double> ll(n_rows), w(n_rows), rate(n_rows),
        pred((size_t)n_rows * std::max(E->nobs, 1));
    rq_batch_out o = one_row_out(st.data(), ll.data(), pred.data(), it.data(), w.data(),
                                 rate.data());
    const int nmax = (int)(sizeof(out->block_size) / sizeof(out->block_size[0]));
    out->n_candidates = 0;
    for (const Cand &c : cands) {
        if (out->n_candidates >= nmax) break;
        E->block = c.block;
        E->block_user = true;
        E->settings.grid_blocks = c.grid;
        double ms = std::numeric_limits<double>::quiet_NaN();
        try {
            evaluate_impl(E, theta, n_rows, n_params, 1, &o, nullptr, nullptr, opts.data());
            // a grid request the device cannot hold falls back to one workgroup
            if (c.grid == 1 || E->last_nblk == c.grid) ms = E->last_ms;
        } catch (const Fail &) {
            // a shape this device cannot launch is skipped
        }
        const int k = out->n_candidates++;
        out->block_size[k] = c.block;
        out->grid_blocks[k] = c.grid;
        out->amg_block_rows[k] = E->settings.amg_block_rows;
        out->kernel_ms[k] = ms;
    }
    const int pick = first_within_slack(out->kernel_ms, out->n_candidates);
    double best = pick >= 0 ? out->kernel_ms[pick] : std::numeric_limits<double>::infinity();
    const int best_block = pick >= 0 ? out->block_size[pick] : saved_block;
    const int best_grid = pick >= 0 ? out->grid_blocks[pick] : saved_grid;
    if (pick >= 0) {
        E->block = best_block;
        E->block_user = true;
        E->settings.grid_blocks = best_grid;
    } else {
        E->block = saved_block;
        E->block_user = saved_user;
        E->settings.grid_blocks = saved_grid;
    }
    // multi-workgroup draws with a multigrid pressure solve: the threshold
    // below which a level runs on one workgroup, at the chosen shape (it
    // changes only which workgroups compute a row, not the row's arithmetic)
    const int saved_rows = E->settings.amg_block_rows;
    if (pick >= 0 && best_grid > 1 && E->mesh_host.amg.levels > 1) {
        // the configured threshold first, then the alternatives
        double rows_ms[4] = {best, 0, 0, 0};
        int rows_of[4] = {saved_rows, 0, 0, 0};
        int nr = 1;
        for (int rows : {256, 512, 2048}) {
            if (out->n_candidates >= nmax || rows == saved_rows) continue;
            E->settings.amg_block_rows = rows;
            double ms = std::numeric_limits<double>::quiet_NaN();
            try {
                evaluate_impl(E, theta, n_rows, n_params, 1, &o, nullptr, nullptr, opts.data());
                if (E->last_nblk == best_grid) ms = E->last_ms;
            } catch (const Fail &) {
            }
            const int k = out->n_candidates++;
            out->block_size[k] = best_block;
            out->grid_blocks[k] = best_grid;
            out->amg_block_rows[k] = rows;
            out->kernel_ms[k] = ms;
            rows_ms[nr] = ms;
            rows_of[nr++] = rows;
        }
        const int r = first_within_slack(rows_ms, nr);
        E->settings.amg_block_rows = r >= 0 ? rows_of[r] : saved_rows;
    }
    out->chosen_block_size = E->block;
    out->chosen_grid_blocks = E->settings.grid_blocks;
    out->chosen_amg_block_rows = E->settings.amg_block_rows;
    return RQ_OK;
}

}  // namespace

// ----------------------------------------------------------------------------
// Serving: registered problems, a persistent grid of workgroups (one slot
// each, as many as the HBM budget and the device's co-resident capacity
// allow), a request ring and a result ring in pinned host memory.
// ----------------------------------------------------------------------------
struct rq_server {
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<rq_engine *> problems;  // owned: validated problem copies
    DevProblem *d_problems = nullptr;
    ServeControl *ctl = nullptr;        // pinned, coherent
    ServeRequest *ring = nullptr;
    ServeOutput *outs = nullptr;
    int ring_size = 0;
    unsigned long long *d_head = nullptr;
    double *d_arenas = nullptr;
    int64_t arena_stride = 0;           // doubles per slot arena
    double *d_amg = nullptr;
    int64_t amg_stride = 0;             // doubles per slot multigrid workspace
    int slots = 0, max_slots = 0;
    int block = kServeBlock;            // threads per serving workgroup
    int64_t budget = 0;
    size_t lds_bytes = 0;
    long long spin_limit = 0;
    bool prepared = false;
    // tickets: host-visible ids; admitted ones map to a device sequence number
    struct Ticket {
        int problem = -1;
        long long seq = -1;          // device sequence, -1 when settled on the host
        bool done = false;
        DevResult res{};
        std::vector<double> pred;
        int status = 0;
        double loglik = 0.0;
        std::string msg;
    };
    std::vector<Ticket> tickets;
    std::vector<long long> slot_seq;   // device sequence occupying each ring slot, -1 free
    std::vector<int64_t> seq_ticket;    // ring slot -> ticket id
    long long next_seq = 0;
    int64_t completed = 0;

    ~rq_server() {
        if (ctl) {
            __atomic_store_n(&ctl->stop, 1, __ATOMIC_RELEASE);
            if (stream) (void)hipStreamSynchronize(stream);
        }
        for (rq_engine *p : problems) delete p;
        if (d_problems) (void)hipFree(d_problems);
        if (d_head) (void)hipFree(d_head);
        if (d_arenas) (void)hipFree(d_arenas);
        if (d_amg) (void)hipFree(d_amg);
        if (ctl) (void)hipHostFree(ctl);
        if (ring) (void)hipHostFree(ring);
        if (outs) (void)hipHostFree(outs);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

bool server_running(rq_server *S) {
    const hipError_t q = hipStreamQuery(S->stream);
    return q == hipErrorNotReady;
}

// Sizes the slot pool from the HBM budget once the problem set is known.
void server_prepare(rq_server *S) {
    if (S->prepared) return;
    if (S->problems.empty()) throw Fail{RQ_EINVAL, "no problem registered"};
    HIPCHECK(hipSetDevice(S->device));
    int maxLds = 0, cus = 0;
    HIPCHECK(hipDeviceGetAttribute(&maxLds, hipDeviceAttributeMaxSharedMemoryPerBlock, S->device));
    HIPCHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, S->device));
    std::vector<DevProblem> table(S->problems.size());
    int64_t stride = 2, amgw = 0;
    size_t lds = 0;
    // 512 threads when a mesh has more than 256 cells (the batch path's
    // default for such meshes); smaller meshes then run the block Krylov
    // path instead of the one-wave path
    S->block = kServeBlock;
    for (const rq_engine *E : S->problems)
        if (E->nc > 256) S->block = kServeBlockLarge;
    for (size_t i = 0; i < S->problems.size(); ++i) {
        const rq_engine *E = S->problems[i];
        const size_t bytes = (size_t)E->lay.stride * sizeof(double);
        table[i].mesh = E->mesh_host;
        table[i].settings = E->settings;
        table[i].lds = bytes + 8192 <= (size_t)maxLds ? 1 : 0;
        if (table[i].lds) lds = std::max(lds, bytes);
        else stride = std::max<int64_t>(stride, E->lay.stride);
        amgw = std::max<int64_t>(amgw, E->mesh_host.amg.workspace);
    }
    stride = (stride + 1) & ~int64_t(1);
    const int64_t slot_bytes = (stride + amgw) * 8;
    int64_t by_budget = S->budget > 0 ? S->budget / slot_bytes : S->max_slots;
    int slots = (int)std::min<int64_t>(S->max_slots, by_budget);
    slots = std::min(slots, cus);  // one co-resident workgroup per CU
    if (slots < 1) throw Fail{RQ_ENOMEM, "HBM budget does not hold one solve slot"};
    HIPCHECK(hipMalloc(&S->d_problems, table.size() * sizeof(DevProblem)));
    HIPCHECK(hipMemcpy(S->d_problems, table.data(), table.size() * sizeof(DevProblem),
                       hipMemcpyHostToDevice));
    HIPCHECK(hipMalloc(&S->d_arenas, (size_t)slots * stride * 8));
    if (amgw > 0) HIPCHECK(hipMalloc(&S->d_amg, (size_t)slots * amgw * 8));
    S->arena_stride = stride;
    S->amg_stride = amgw;
    S->slots = slots;
    S->lds_bytes = lds;
    S->prepared = true;
}

void server_ensure_running(rq_server *S) {
    if (server_running(S)) return;
    HIPCHECK(hipSetDevice(S->device));
    HIPCHECK(launch_serve(S->block, S->slots, S->d_problems, S->ctl, S->ring, S->outs, S->ring_size,
                          S->d_head, S->d_arenas, S->arena_stride, S->d_amg, S->amg_stride,
                          S->spin_limit, S->lds_bytes, S->stream));
}

// Moves finished device results of the ring into their tickets.
void server_collect(rq_server *S) {
    for (int slot = 0; slot < S->ring_size; ++slot) {
        const long long seq = S->slot_seq[slot];
        if (seq < 0) continue;
        const long long done = __atomic_load_n(&S->outs[slot].done, __ATOMIC_ACQUIRE);
        if (done != seq + 1) continue;
        rq_server::Ticket &t = S->tickets[S->seq_ticket[slot]];
        const rq_engine *E = S->problems[t.problem];
        t.res = S->outs[slot].res;
        t.pred.assign(S->outs[slot].predictions, S->outs[slot].predictions + E->nobs);
        const RowFinal fin = finalize_row(E, t.res, t.pred.data());
        t.status = fin.status;
        t.loglik = fin.loglik;
        t.msg = fin.msg ? fin.msg : "";
        if (!fin.keep_predictions)
            std::fill(t.pred.begin(), t.pred.end(), std::numeric_limits<double>::quiet_NaN());
        t.done = true;
        S->slot_seq[slot] = -1;
        ++S->completed;
    }
}

void fill_row(const rq_server::Ticket &t, const rq_engine *E, rq_serve_row *row) {
    row->status = t.status;
    row->loglik = t.loglik;
    if (row->predictions)
        for (int j = 0; j < E->nobs; ++j) row->predictions[j] = t.pred[j];
    const bool solved = t.seq >= 0;
    row->iterations = solved ? t.res.iters : 0;
    row->final_worst_measure = solved ? t.res.worst : 0.0;
    row->contraction_rate_u = solved ? t.res.rate : 0.0;
    row->stalled = solved ? (uint8_t)t.res.stalled : 0;
    row->converged = t.status == ST_CONVERGED;
    for (int e = 0; e < 5; ++e) row->krylov_iterations[e] = solved ? t.res.krylov[e] : 0;
    std::snprintf(row->error, sizeof row->error, "%s", t.msg.c_str());
}

}  // namespace

extern "C" {

rq_engine *rq_engine_create(const rq_mesh_desc *mesh, const rq_param_desc *params,
                            const rq_obs_desc *obs, const rq_bc_desc *bcs, double nu,
                            const rq_settings *settings, const double init[5], char *err,
                            size_t errcap) {
    try {
        rq_engine *e = create_impl(mesh, params, obs, bcs, nu, settings, init);
        set_err(err, errcap, "");
        return e;
    } catch (const Fail &f) {
        set_err(err, errcap, f.msg);
    } catch (const std::bad_alloc &) {
        set_err(err, errcap, "host allocation failed");
    } catch (const std::exception &e) {
        set_err(err, errcap, e.what());
    }
    return nullptr;
}

int rq_engine_evaluate(rq_engine *engine, const double *theta, int64_t n_rows, int32_t n_params,
                       int32_t workers, rq_batch_out *out, char *err, size_t errcap) {
    try {
        int rc = evaluate_impl(engine, theta, n_rows, n_params, workers, out);
        set_err(err, errcap, "");
        return rc;
    } catch (const Fail &f) {
        set_err(err, errcap, f.msg);
        return f.code;
    } catch (const std::bad_alloc &) {
        set_err(err, errcap, "host allocation failed");
        return RQ_ENOMEM;
    } catch (const std::exception &e) {
        set_err(err, errcap, e.what());
        return RQ_EDEVICE;
    }
}

int rq_engine_autotune(rq_engine *engine, const double *theta, int64_t n_rows, int32_t n_params,
                       int32_t iterations, rq_tune_out *out, char *err, size_t errcap) {
    return guarded(err, errcap, [&]() {
        return autotune_impl(engine, theta, n_rows, n_params, iterations, out);
    });
}

int rq_engine_set_graph_replay(rq_engine *e, int32_t enable, char *err, size_t errcap) {
    try {
        if (!e) throw Fail{RQ_EINVAL, "null engine"};
        if (enable != 0 && enable != 1) throw Fail{RQ_EINVAL, "enable must be 0 or 1"};
        std::lock_guard<std::recursive_mutex> lock(e->mu);
        HIPCHECK(hipSetDevice(e->device));
        e->graph_on = enable != 0;
        if (!e->graph_on) e->drop_graph();
        set_err(err, errcap, "");
        return RQ_OK;
    } catch (const Fail &f) {
        set_err(err, errcap, f.msg);
        return f.code;
    }
}

int64_t rq_engine_graph_launches(const rq_engine *e) { return e ? e->graph_launches : -1; }

int32_t rq_engine_amg_levels(const rq_engine *e, int32_t *rows, int64_t *entries, int32_t cap) {
    if (!e) return -1;
    const AmgMesh &G = e->mesh_host.amg;
    for (int l = 0; l < G.levels && l < cap; ++l) {
        if (rows) rows[l] = G.lev[l].n;
        if (entries) entries[l] = G.lev[l].nnz;
    }
    return G.levels;
}

// ----------------------------------------------------------------------------
// Allocation-free evaluation (no counterpart in the reference): the kernels
// of rq_engine_evaluate, with every per-call device buffer (slot table,
// result rows, predictions, grid-barrier state, HBM arenas, multigrid
// workspace) carved out of the current GPU's fixed scratch region, and the
// host side staged through small stack buffers. Errors are return codes
// (nothing is thrown), so no path of the call allocates.
// ----------------------------------------------------------------------------
static size_t round256(size_t b) { return (b + 255) & ~(size_t)255; }

static void put_err(char *err, size_t cap, const char *fmt, ...) {
    if (!err || cap == 0) return;
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(err, cap, fmt, ap);
    va_end(ap);
}

static_assert(sizeof(GridState) % sizeof(double) == 0, "grid state is cleared as doubles");
static_assert(kNoAllocScratchBytes == RQ_NOALLOC_SCRATCH_BYTES, "one scratch size");

int rq_engine_evaluate_noalloc(rq_engine *E, const double *theta, int64_t n_rows, int32_t n_params,
                               int32_t workers, rq_batch_out *out, char *err, size_t errcap) {
#define RQ_NA_FAIL(code, ...)                     \
    do {                                          \
        put_err(err, errcap, __VA_ARGS__);        \
        return code;                              \
    } while (0)
#define RQ_NA_HIP(expr)                                                                    \
    do {                                                                                   \
        const hipError_t e_ = (expr);                                                      \
        if (e_ != hipSuccess) RQ_NA_FAIL(RQ_EDEVICE, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)
    if (!E || !out || !out->status) RQ_NA_FAIL(RQ_EINVAL, "null argument");
    std::lock_guard<std::recursive_mutex> lock(E->mu);
    if (workers < 1) RQ_NA_FAIL(RQ_EINVAL, "workers must be positive");
    if (n_params != E->n_active)
        RQ_NA_FAIL(RQ_EINVAL, "batch rows must all have n_parameters entries");
    if (n_rows < 0) RQ_NA_FAIL(RQ_EINVAL, "row count must be nonnegative");
    if (n_rows > 0 && !theta) RQ_NA_FAIL(RQ_EINVAL, "null coefficient rows");
    if (out->fields) RQ_NA_FAIL(RQ_EINVAL, "the allocation-free evaluation returns no fields");
    const int nobs = E->nobs, po = std::max(nobs, 1);
    if (n_rows > 0 && nobs > 0 && !out->predictions)
        RQ_NA_FAIL(RQ_EINVAL, "the allocation-free evaluation needs the predictions array");
    // the region's plan for n_rows rows (every row admitted), checked before
    // anything is written
    const size_t cap = kNoAllocScratchBytes;
    if (n_rows > (int64_t)(cap / sizeof(DevSlot)))
        RQ_NA_FAIL(RQ_EINVAL,
                   "the allocation-free evaluation takes at most %zu bytes of device scratch; "
                   "%lld rows exceed it", cap, (long long)n_rows);
    const bool arenas = !E->lds || E->settings.grid_blocks > 1;
    const int64_t amgw = E->mesh_host.amg.workspace;
    const size_t nr = (size_t)n_rows;
    size_t off = 0;
    const size_t o_slots = off;
    off += round256(nr * sizeof(DevSlot));
    const size_t o_res = off;
    off += round256(nr * sizeof(DevResult));
    const size_t o_pred = off;
    off += round256(nr * po * sizeof(double));
    const size_t o_grid = off;
    off += arenas ? round256(nr * sizeof(GridState)) : 0;
    const size_t o_arena = off;
    off += arenas ? round256(nr * (size_t)E->lay.stride * sizeof(double)) : 0;
    const size_t o_amg = off;
    off += amgw > 0 ? round256(nr * (size_t)amgw * sizeof(double)) : 0;
    if (off > cap)
        RQ_NA_FAIL(RQ_EINVAL,
                   "the allocation-free evaluation of %lld rows needs %zu bytes of device "
                   "scratch; its fixed region holds %zu bytes", (long long)n_rows, off, cap);
    const double nan = std::numeric_limits<double>::quiet_NaN();
    // per-row outputs and the host admission (ForwardModel.cpp:45-49);
    // admitted rows are marked -1 until their result arrives
    int ns = 0;
    for (int64_t r = 0; r < n_rows; ++r) {
        const double *row = theta + r * n_params;
        if (out->row_errors && out->row_error_stride) out->row_errors[r * out->row_error_stride] = 0;
        if (out->predictions)
            for (int i = 0; i < nobs; ++i) out->predictions[r * nobs + i] = nan;
        if (out->iterations) out->iterations[r] = 0;
        if (out->final_worst_measure) out->final_worst_measure[r] = 0.0;
        if (out->contraction_rate_u) out->contraction_rate_u[r] = 0.0;
        if (out->stalled) out->stalled[r] = 0;
        if (out->converged) out->converged[r] = 0;
        if (out->krylov_iterations)
            for (int e = 0; e < 5; ++e) out->krylov_iterations[r * 5 + e] = 0;
        if (out->phase_cycles)
            for (int e = 0; e < kPhases; ++e) out->phase_cycles[r * kPhases + e] = 0;
        double coeffs[11];
        const char *msg = nullptr;
        const int bad = admit_row(E, row, coeffs, msg);
        if (bad) {
            out->status[r] = bad;
            if (out->loglik) out->loglik[r] = bad == ST_INVALID ? -1e30 : -1e6;
            if (msg) write_row_error(out, r, msg);
        } else {
            out->status[r] = -1;
            ++ns;
        }
    }
    if (ns == 0) {
        put_err(err, errcap, "%s", "");
        return RQ_OK;
    }
    RQ_NA_HIP(hipSetDevice(E->device));
    void *base = nullptr;
    size_t have = 0;
    RQ_NA_HIP(noalloc_scratch(&base, &have));
    if (off > have) RQ_NA_FAIL(RQ_EINVAL, "device scratch region smaller than planned");
    // every call on this GPU shares the region: one at a time, from the first
    // write to the last read
    std::lock_guard<std::mutex> region(noalloc_lock(E->device));
    unsigned char *b = static_cast<unsigned char *>(base);
    DevSlot *d_slots = reinterpret_cast<DevSlot *>(b + o_slots);
    DevResult *d_res = reinterpret_cast<DevResult *>(b + o_res);
    double *d_pred = reinterpret_cast<double *>(b + o_pred);
    GridState *d_grid = reinterpret_cast<GridState *>(b + o_grid);
    double *d_arena = reinterpret_cast<double *>(b + o_arena);
    double *d_amg = reinterpret_cast<double *>(b + o_amg);
    const size_t grid_lds = grid_lds_bytes(E);
    const int nblk = choose_grid_blocks(E, ns, grid_lds);
    // the slot table, uploaded from the stack a few rows at a time; a row's
    // result and predictions land at its own row index
    constexpr int kStage = 32;
    DevSlot stage[kStage];
    int placed = 0, k = 0;
    for (int64_t r = 0; r < n_rows; ++r) {
        if (out->status[r] != -1) continue;
        DevSlot &sl = stage[k];
        sl = DevSlot{};
        sl.mesh = E->d_mesh;
        const char *msg = nullptr;
        (void)admit_row(E, theta + r * n_params, sl.coeffs, msg);
        sl.row = (int)r;
        if (!E->lds || nblk > 1) sl.arena = d_arena + (size_t)placed * E->lay.stride;
        if (amgw > 0) sl.amg = d_amg + (size_t)placed * amgw;
        ++placed;
        if (++k == kStage || placed == ns) {
            RQ_NA_HIP(hipMemcpyAsync(d_slots + (placed - k), stage, k * sizeof(DevSlot),
                                     hipMemcpyHostToDevice, E->stream));
            RQ_NA_HIP(hipStreamSynchronize(E->stream));
            k = 0;
        }
    }
    // predictions start as NaN (rows the host settled keep them); the
    // barrier counters of a multi-workgroup launch start at zero
    RQ_NA_HIP(launch_fill(d_pred, (int64_t)(nr * po), nan, E->stream));
    if (nblk > 1)
        RQ_NA_HIP(launch_fill(reinterpret_cast<double *>(d_grid),
                              (int64_t)(ns * sizeof(GridState) / sizeof(double)), 0.0, E->stream));
    RQ_NA_HIP(hipEventRecord(E->ev0, E->stream));
    if (nblk > 1)
        RQ_NA_HIP(launch_grid_solve(d_slots, ns, nblk, E->settings, d_res, d_pred, po, nullptr, 0,
                                    d_grid, grid_lds, E->stream));
    else
        RQ_NA_HIP(launch_simple_solve(E->block, E->lds, d_slots, ns, E->settings, d_res, d_pred,
                                      po, nullptr, 0, E->lds_bytes, E->stream));
    RQ_NA_HIP(hipEventRecord(E->ev1, E->stream));
    if (nobs > 0)
        RQ_NA_HIP(hipMemcpyAsync(out->predictions, d_pred, nr * nobs * sizeof(double),
                                 hipMemcpyDeviceToHost, E->stream));
    RQ_NA_HIP(hipStreamSynchronize(E->stream));
    float ms = 0.f;
    RQ_NA_HIP(hipEventElapsedTime(&ms, E->ev0, E->ev1));
    E->last_ms = ms;
    E->last_nblk = nblk;
    DevResult rstage[kStage];
    for (int64_t r0 = 0; r0 < n_rows; r0 += kStage) {
        const int cnt = (int)std::min<int64_t>(kStage, n_rows - r0);
        bool any = false;
        for (int j = 0; j < cnt; ++j) any |= out->status[r0 + j] == -1;
        if (!any) continue;
        RQ_NA_HIP(hipMemcpyAsync(rstage, d_res + r0, cnt * sizeof(DevResult),
                                 hipMemcpyDeviceToHost, E->stream));
        RQ_NA_HIP(hipStreamSynchronize(E->stream));
        for (int j = 0; j < cnt; ++j) {
            const int64_t r = r0 + j;
            if (out->status[r] != -1) continue;
            const DevResult &d = rstage[j];
            const RowFinal fin =
                finalize_row(E, d, nobs > 0 ? out->predictions + r * nobs : nullptr);
            if (!fin.keep_predictions && out->predictions)
                for (int i = 0; i < nobs; ++i) out->predictions[r * nobs + i] = nan;
            out->status[r] = fin.status;
            if (out->loglik) out->loglik[r] = fin.loglik;
            if (out->iterations) out->iterations[r] = d.iters;
            if (out->final_worst_measure) out->final_worst_measure[r] = d.worst;
            if (out->contraction_rate_u) out->contraction_rate_u[r] = d.rate;
            if (out->stalled) out->stalled[r] = (uint8_t)d.stalled;
            if (out->converged) out->converged[r] = fin.status == ST_CONVERGED;
            if (out->krylov_iterations)
                for (int e = 0; e < 5; ++e) out->krylov_iterations[r * 5 + e] = d.krylov[e];
            if (out->phase_cycles)
                for (int e = 0; e < kPhases; ++e) out->phase_cycles[r * kPhases + e] = d.cycles[e];
            if (fin.msg) write_row_error(out, r, fin.msg);
        }
    }
    put_err(err, errcap, "%s", "");
    return RQ_OK;
#undef RQ_NA_HIP
#undef RQ_NA_FAIL
}

int32_t rq_engine_n_observations(const rq_engine *e) { return e ? e->nobs : -1; }
int32_t rq_engine_n_parameters(const rq_engine *e) { return e ? e->n_active : -1; }
int32_t rq_engine_n_cells(const rq_engine *e) { return e ? e->nc : -1; }
int32_t rq_engine_n_faces(const rq_engine *e) { return e ? e->nf : -1; }

int rq_engine_last_timing(const rq_engine *e, double *kernel_ms, int32_t *block_size,
                          int32_t *arena_in_lds, int64_t *arena_bytes, int32_t *grid_blocks) {
    if (!e) return RQ_EINVAL;
    std::lock_guard<std::recursive_mutex> lock(e->mu);
    if (kernel_ms) *kernel_ms = e->last_ms;
    if (block_size) *block_size = e->last_nblk > 1 ? kGridBlock : e->block;
    if (arena_in_lds) *arena_in_lds = (e->lds && e->last_nblk == 1) ? 1 : 0;
    if (grid_blocks) *grid_blocks = e->last_nblk;
    if (arena_bytes) *arena_bytes = e->lay.stride * (int64_t)sizeof(double);
    return RQ_OK;
}

void rq_engine_destroy(rq_engine *e) { delete e; }

const char *rq_build_info(void) {
    return "rans_amd: gfx950 HIP, FP64 assembly and convergence, -O3 -ffp-contract=off"
#ifdef __clang_version__
           " clang " __clang_version__
#endif
        ;
}


rq_server *rq_server_create(const rq_serve_options *opt, char *err, size_t errcap) {
    rq_server *S = nullptr;
    try {
        if (!opt) throw Fail{RQ_EINVAL, "null argument"};
        if (opt->max_slots < 1) throw Fail{RQ_EINVAL, "max_slots must be positive"};
        if (opt->ring_size < 1 || opt->ring_size > (1 << 20))
            throw Fail{RQ_EINVAL, "ring_size must be in [1, 2^20]"};
        if (opt->hbm_budget_bytes < 0) throw Fail{RQ_EINVAL, "HBM budget must be nonnegative"};
        if (!(opt->idle_exit_seconds > 0.0) || opt->idle_exit_seconds > 3600.0)
            throw Fail{RQ_EINVAL, "idle_exit_seconds must be in (0, 3600]"};
        S = new rq_server();
        HIPCHECK(hipGetDevice(&S->device));
        S->max_slots = opt->max_slots;
        S->budget = opt->hbm_budget_bytes;
        S->ring_size = opt->ring_size;
        // one idle poll sleeps about 100 x 64 shader clocks (~2.7 us at 2.4 GHz)
        S->spin_limit = (long long)(opt->idle_exit_seconds / 2.7e-6) + 1;
        HIPCHECK(hipHostMalloc((void **)&S->ctl, sizeof(ServeControl),
                               hipHostMallocCoherent | hipHostMallocMapped));
        HIPCHECK(hipHostMalloc((void **)&S->ring, sizeof(ServeRequest) * S->ring_size,
                               hipHostMallocCoherent | hipHostMallocMapped));
        HIPCHECK(hipHostMalloc((void **)&S->outs, sizeof(ServeOutput) * S->ring_size,
                               hipHostMallocCoherent | hipHostMallocMapped));
        std::memset(S->ctl, 0, sizeof(ServeControl));
        std::memset(S->ring, 0, sizeof(ServeRequest) * S->ring_size);
        std::memset(S->outs, 0, sizeof(ServeOutput) * S->ring_size);
        // [0] tickets claimed, [1] the device-side mirror of the ring tail
        HIPCHECK(hipMalloc(&S->d_head, 2 * sizeof(unsigned long long)));
        HIPCHECK(hipMemset(S->d_head, 0, 2 * sizeof(unsigned long long)));
        HIPCHECK(hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking));
        S->slot_seq.assign(S->ring_size, -1);
        S->seq_ticket.assign(S->ring_size, -1);
        set_err(err, errcap, "");
        return S;
    } catch (const Fail &f) {
        set_err(err, errcap, f.msg);
    } catch (const std::exception &ex) {
        set_err(err, errcap, ex.what());
    }
    delete S;
    return nullptr;
}

int32_t rq_server_add_problem(rq_server *S, const rq_mesh_desc *mesh, const rq_param_desc *params,
                              const rq_obs_desc *obs, const rq_bc_desc *bcs, double nu,
                              const rq_settings *settings, const double init[5], char *err,
                              size_t errcap) {
    try {
        if (!S) throw Fail{RQ_EINVAL, "null server"};
        if (S->prepared)
            throw Fail{RQ_EINVAL, "problems are registered before the first submit"};
        if (settings && settings->mixed_precision)
            throw Fail{RQ_EINVAL, "served problems run in FP64"};
        if (obs && obs->n_obs > kMaxServeObs)
            throw Fail{RQ_EINVAL, "a served problem has at most 64 observations"};
        rq_engine *E = create_impl(mesh, params, obs, bcs, nu, settings, init);
        S->problems.push_back(E);
        set_err(err, errcap, "");
        return (int32_t)S->problems.size() - 1;
    } catch (const Fail &f) {
        set_err(err, errcap, f.msg);
    } catch (const std::bad_alloc &) {
        set_err(err, errcap, "host allocation failed");
    } catch (const std::exception &ex) {
        set_err(err, errcap, ex.what());
    }
    return -1;
}

int64_t rq_server_submit(rq_server *S, int32_t problem, const double *theta, int32_t n_params,
                         char *err, size_t errcap) {
    try {
        if (!S || !theta) throw Fail{RQ_EINVAL, "null argument"};
        if (problem < 0 || problem >= (int32_t)S->problems.size())
            throw Fail{RQ_EINVAL, "unknown problem"};
        const rq_engine *E = S->problems[problem];
        if (n_params != E->n_active)
            throw Fail{RQ_EINVAL, "theta must have n_parameters entries"};
        server_prepare(S);
        rq_server::Ticket t;
        t.problem = problem;
        double coeffs[11];
        const char *msg = nullptr;
        const int bad = admit_row(E, theta, coeffs, msg);
        if (bad) {  // settled on the host, like a batch row that never reaches the device
            t.done = true;
            t.status = bad;
            t.loglik = bad == ST_INVALID ? -1e30 : -1e6;
            t.msg = msg ? msg : "";
            t.pred.assign(E->nobs, std::numeric_limits<double>::quiet_NaN());
            S->tickets.push_back(std::move(t));
            set_err(err, errcap, "");
            return (int64_t)S->tickets.size() - 1;
        }
        server_collect(S);
        const long long seq = S->next_seq;
        const int slot = (int)(seq % S->ring_size);
        if (S->slot_seq[slot] >= 0) throw Fail{RQ_EAGAIN, "request ring is full"};
        ServeRequest &rq = S->ring[slot];
        rq.ticket = seq;
        rq.problem = problem;
        for (int i = 0; i < 11; ++i) rq.coeffs[i] = coeffs[i];
        S->outs[slot].done = 0;
        t.seq = seq;
        S->tickets.push_back(std::move(t));
        S->slot_seq[slot] = seq;
        S->seq_ticket[slot] = (int64_t)S->tickets.size() - 1;
        S->next_seq = seq + 1;
        __atomic_store_n(&S->ctl->tail, seq + 1, __ATOMIC_RELEASE);
        server_ensure_running(S);
        set_err(err, errcap, "");
        return (int64_t)S->tickets.size() - 1;
    } catch (const Fail &f) {
        set_err(err, errcap, f.msg);
        return f.code == RQ_EAGAIN ? -2 : -1;
    } catch (const std::exception &ex) {
        set_err(err, errcap, ex.what());
    }
    return -1;
}

int rq_server_poll(rq_server *S, int64_t ticket, rq_serve_row *row, int32_t *done, char *err,
                   size_t errcap) {
    try {
        if (!S || !done) throw Fail{RQ_EINVAL, "null argument"};
        if (ticket < 0 || ticket >= (int64_t)S->tickets.size())
            throw Fail{RQ_EINVAL, "unknown ticket"};
        *done = 0;
        rq_server::Ticket &t = S->tickets[ticket];
        if (!t.done) {
            server_collect(S);
            if (!t.done) server_ensure_running(S);
        }
        if (t.done) {
            *done = 1;
            if (row) fill_row(t, S->problems[t.problem], row);
        }
        set_err(err, errcap, "");
        return RQ_OK;
    } catch (const Fail &f) {
        set_err(err, errcap, f.msg);
        return f.code;
    } catch (const std::exception &ex) {
        set_err(err, errcap, ex.what());
    }
    return RQ_EDEVICE;
}

int rq_server_stats(const rq_server *S, int32_t *slots, int64_t *slot_arena_bytes,
                    int64_t *submitted, int64_t *completed) {
    if (!S) return RQ_EINVAL;
    if (slots) *slots = S->slots;
    if (slot_arena_bytes) *slot_arena_bytes = S->arena_stride * 8;
    if (submitted) *submitted = S->next_seq;
    if (completed) *completed = S->completed;
    return RQ_OK;
}

void rq_server_destroy(rq_server *S) { delete S; }

// ---------------------------------------------------------------- model
rq_model *rq_model_create(const rq_mesh_desc *mesh, const rq_param_desc *params,
                          const rq_obs_desc *obs, const rq_bc_desc *bcs, double nu,
                          const rq_settings *settings, const double init[5],
                          const rq_cache_options *cache, char *err, size_t errcap) {
    try {
        rq::WarmCache c = cache ? rq::WarmCache(cache->max_size, cache->threshold,
                                                cache->byte_budget < 0 ? 0 : (size_t)cache->byte_budget)
                                : rq::WarmCache();
        if (cache && cache->byte_budget < 0)
            throw Fail{RQ_EINVAL, "Invalid warm-start cache limits"};
        auto *M = new rq_model(c);
        try {
            M->E = create_impl(mesh, params, obs, bcs, nu, settings, init);
        } catch (...) {
            delete M;
            throw;
        }
        M->state_n = state_doubles(M->E->nc, M->E->nb);
        set_err(err, errcap, "");
        return M;
    } catch (const Fail &f) {
        set_err(err, errcap, f.msg);
    } catch (const std::invalid_argument &e) {
        set_err(err, errcap, e.what());
    } catch (const std::bad_alloc &) {
        set_err(err, errcap, "host allocation failed");
    } catch (const std::exception &e) {
        set_err(err, errcap, e.what());
    }
    return nullptr;
}

static int model_call(rq_model *m, const double *theta, int64_t n_rows, int32_t n_params,
                      rq_batch_out *out, uint8_t *used, bool batch, char *err, size_t errcap) {
    try {
        int rc = model_evaluate_impl(m, theta, n_rows, n_params, out, used, batch);
        set_err(err, errcap, "");
        return rc;
    } catch (const Fail &f) {
        set_err(err, errcap, f.msg);
        return f.code;
    } catch (const std::bad_alloc &) {
        set_err(err, errcap, "host allocation failed");
        return RQ_ENOMEM;
    } catch (const std::exception &e) {
        set_err(err, errcap, e.what());
        return RQ_EDEVICE;
    }
}

int rq_model_evaluate(rq_model *m, const double *theta, int64_t n_rows, int32_t n_params,
                      rq_batch_out *out, uint8_t *used_warm_start, char *err, size_t errcap) {
    return model_call(m, theta, n_rows, n_params, out, used_warm_start, false, err, errcap);
}

int rq_model_evaluate_batch(rq_model *m, const double *theta, int64_t n_rows, int32_t n_params,
                            rq_batch_out *out, uint8_t *used_warm_start, char *err,
                            size_t errcap) {
    return model_call(m, theta, n_rows, n_params, out, used_warm_start, true, err, errcap);
}

const rq_engine *rq_model_engine(const rq_model *m) { return m ? m->E : nullptr; }
int32_t rq_model_cache_size(const rq_model *m) { return m ? m->cache.size() : -1; }
int64_t rq_model_cache_retained_bytes(const rq_model *m) {
    return m ? (int64_t)m->cache.retained_bytes() : -1;
}
int64_t rq_model_cache_byte_budget(const rq_model *m) {
    return m ? (int64_t)m->cache.byte_budget() : -1;
}
int rq_model_cache_set_byte_budget(rq_model *m, int64_t bytes) {
    if (!m || bytes < 0) return RQ_EINVAL;
    std::vector<void *> freed;
    m->cache.set_byte_budget((size_t)bytes, freed);
    m->give_back(freed);
    return RQ_OK;
}
int rq_model_cache_clear(rq_model *m) {
    if (!m) return RQ_EINVAL;
    std::vector<void *> freed;
    m->cache.clear(freed);
    m->give_back(freed);
    return RQ_OK;
}
void rq_model_destroy(rq_model *m) { delete m; }

// ------------------------------------------- cache bookkeeping without a GPU
struct rq_warm_cache {
    rq::WarmCache c;
};

rq_warm_cache *rq_warm_cache_create(int32_t max_size, double threshold, int64_t byte_budget,
                                    char *err, size_t errcap) {
    try {
        if (byte_budget < 0) throw std::invalid_argument("Invalid warm-start cache limits");
        auto *w = new rq_warm_cache{rq::WarmCache(max_size, threshold, (size_t)byte_budget)};
        set_err(err, errcap, "");
        return w;
    } catch (const std::exception &e) {
        set_err(err, errcap, e.what());
        return nullptr;
    }
}
int32_t rq_warm_cache_store(rq_warm_cache *w, const double *theta, int32_t n, int64_t tag,
                            int64_t bytes) {
    if (!w || n < 0 || (n > 0 && !theta) || bytes < 0) return -1;
    std::vector<void *> freed;
    return w->c.store(theta, (size_t)n, reinterpret_cast<void *>(tag), (size_t)bytes, freed) ? 1
                                                                                           : 0;
}
int64_t rq_warm_cache_nearest(const rq_warm_cache *w, const double *theta, int32_t n) {
    if (!w || n < 0 || (n > 0 && !theta)) return -1;
    const rq::WarmCache::Entry *e = w->c.nearest(theta, (size_t)n);
    return e ? reinterpret_cast<int64_t>(e->payload) : -1;
}
int32_t rq_warm_cache_size(const rq_warm_cache *w) { return w ? w->c.size() : -1; }
int64_t rq_warm_cache_retained_bytes(const rq_warm_cache *w) {
    return w ? (int64_t)w->c.retained_bytes() : -1;
}
int rq_warm_cache_set_byte_budget(rq_warm_cache *w, int64_t bytes) {
    if (!w || bytes < 0) return RQ_EINVAL;
    std::vector<void *> freed;
    w->c.set_byte_budget((size_t)bytes, freed);
    return RQ_OK;
}
int rq_warm_cache_clear(rq_warm_cache *w) {
    if (!w) return RQ_EINVAL;
    std::vector<void *> freed;
    w->c.clear(freed);
    return RQ_OK;
}
int64_t rq_warm_cache_entry_bytes(int32_t n_theta, int32_t n_cells, int32_t n_boundary) {
    return (int64_t)rq::WarmCache::entry_bytes((size_t)n_theta, n_cells, n_boundary);
}
void rq_warm_cache_destroy(rq_warm_cache *w) { delete w; }

// ---------------------------------------------------------------- sensitivity
rq_sensitivity *rq_sensitivity_create(const rq_mesh_desc *mesh, const rq_obs_desc *obs,
                                      const rq_bc_desc *bcs, double nu,
                                      const rq_settings *settings, const double init[5],
                                      char *err, size_t errcap) {
    rq_sensitivity *S = nullptr;
    const int rc = guarded(err, errcap, [&]() {
        int32_t idx[11];
        double base[11], lo[11], hi[11];
        for (int j = 0; j < 11; ++j) {
            idx[j] = j;
            base[j] = kAll11Lower[j];
            lo[j] = 0.0;  // admission of perturbed rows is decided before the launch
            hi[j] = std::numeric_limits<double>::infinity();
        }
        rq_param_desc P{11, idx, base, lo, hi};
        auto *s = new rq_sensitivity();
        try {
            s->E = create_impl(mesh, &P, obs, bcs, nu, settings, init);
        } catch (...) {
            delete s;
            throw;
        }
        S = s;
        return RQ_OK;
    });
    return rc == RQ_OK ? S : nullptr;
}

int rq_sensitivity_solve_state(rq_sensitivity *s, const double theta11[11], int32_t *status,
                               char *err, size_t errcap) {
    return guarded(err, errcap, [&]() { return solve_state_impl(s, theta11, status); });
}

int32_t rq_sensitivity_has_state(const rq_sensitivity *s) { return s && s->has_state ? 1 : 0; }

int rq_sensitivity_eta_jacobian_warm_fd(rq_sensitivity *s, const double theta11[11],
                                        double h_rel, double h_floor, int32_t warm_max_iter,
                                        double warm_tol, const double *absolute_steps,
                                        rq_warm_fd_out *out, char *err, size_t errcap) {
    return guarded(err, errcap, [&]() {
        return warm_fd_impl(s, theta11, h_rel, h_floor, warm_max_iter, warm_tol, absolute_steps,
                            out);
    });
}

const rq_engine *rq_sensitivity_engine(const rq_sensitivity *s) { return s ? s->E : nullptr; }
void rq_sensitivity_destroy(rq_sensitivity *s) { delete s; }
}  // extern "C"
