// Test-infrastructure only: a driver that runs the reference's own forward
// model (compiled from its sources where they lie) on a case described in a
// small text file and prints one JSON object per coefficient row. It is linked
// only into oracle/_ref/ref_harness and is never part of the product library.
//
// The driver follows the reference's golden recorder pattern
// (tests/tools/record_goldens.cpp:152-165 runSimple, :97-136 field norms) and
// the forward-model path (rans-uq/src/ForwardModel.cpp:35-137).
//
// Case file grammar, one directive per line ('#' starts a comment):
//   mesh_file <path.msh>                       MSH2 mesh (core/src/Mesh.cpp:817-982)
//   mesh_periodic_channel nx ny Lx h ReTau yplus
//   mesh_curved nNodes x0 yb0 x1 yb1 ... yTop ny Re yplus
//   mesh_channel nx ny Lx Ly                   inlet/outlet channel (makeChannel2D)
//   mesh_channel_graded nx ny Lx Ly Re yplus   the same, wall-graded (makeChannel2D overload)
//   mesh_plate nx ny Lx Ly Re yplus            one-sided graded plate mesh (makePlate2D)
//   mesh_bfs nxu nxd nyu nyd Lu Ld h H Re yplus backward-facing step (makeBackwardFacingStep2D)
//   patch_type <name> <type>                   Mesh::setPatchType
//   save_mesh <path.msh>
//   nu <value>
//   bcs_channel Uin kIn omIn
//   bcs_couette Uwall kIn omIn                 moving top wall (couetteDefaults)
//   bcs_plate Uinf kIn omIn                    free-stream top (flatPlateDefaults)
//   set <key> <value>                          SolverSettings field
//   init Ux Uy Uz p k omega
//   obs_u x y comp observed sigma
//   obs_cf patch x y observed sigma refVel
//   obs_p x y observed sigma refVel refP refRho
//   obs_drag patch observed sigma refArea refVel drag coefficient
//   obs_sep patch observed sigma                separation point
//   obs_reatt patch observed sigma              reattachment length
//   params idx...
//   theta v...                                 one coefficient row (repeatable)
//   dump <prefix>                              binary field dump per row
//   warm_cache <bytes>                         keep ForwardModel's warm-start cache with
//                                              this byte budget (default: off)
#include "BoundaryCondition.hpp"
#include "FlowFields.hpp"
#include "ForwardModel.hpp"
#include "InferenceParameters.hpp"
#include "Mesh.hpp"
#include "ObservationOperator.hpp"
#include "SIMPLESolver.hpp"

#include <chrono>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

namespace {

void setKey(SolverSettings &s, const std::string &k, const std::string &v) {
    auto d = [&] { return std::stod(v); };
    auto i = [&] { return std::stoi(v); };
    if (k == "max_iterations") s.maxIterations = i();
    else if (k == "convergence_tol") s.convergenceTol = d();
    else if (k == "divergence_limit") s.divergenceLimit = d();
    else if (k == "alpha_u") s.alphaU = d();
    else if (k == "alpha_p") s.alphaP = d();
    else if (k == "alpha_k") s.alphaK = d();
    else if (k == "alpha_omega") s.alphaOmega = d();
    else if (k == "inner_iterations") s.innerIterations = i();
    else if (k == "inner_tolerance") s.innerTolerance = d();
    else if (k == "turb_start_iter") s.turbStartIter = i();
    else if (k == "turb_update_interval") s.turbUpdateInterval = i();
    else if (k == "nut_floor_iters") s.nuTFloorIters = i();
    else if (k == "convergence_window") s.convergenceWindow = i();
    else if (k == "monitor_interval") s.monitorInterval = i();
    else if (k == "body_force_x") s.bodyForce.x = d();
    else if (k == "body_force_y") s.bodyForce.y = d();
    else if (k == "pressure_solver") s.pressureSolver = v;
    else if (k == "momentum_solver") s.momentumSolver = v;
    else if (k == "turbulence_solver") s.turbulenceSolver = v;
    else if (k == "momentum_convection") s.momentumConvection = v;
    else if (k == "turbulence_convection") s.turbulenceConvection = v;
    else if (k == "k_min") s.kMin = d();
    else if (k == "omega_min") s.omegaMin = d();
    else if (k == "k_max") s.kMax = d();
    else if (k == "sst_variant") s.sstVariant = i();
    else throw std::runtime_error("unknown setting " + k);
}

std::string num(double v) {
    if (std::isnan(v)) return "NaN";
    if (std::isinf(v)) return v > 0 ? "Infinity" : "-Infinity";
    char buf[40];
    std::snprintf(buf, sizeof buf, "%.17g", v);
    return buf;
}

void dumpFields(const std::string &path, const Mesh &m, const FlowFields &f,
                const std::vector<double> &flux) {
    std::ofstream o(path, std::ios::binary);
    const int nc = m.nCells(), nb = m.nFaces() - m.nInternalFaces(), nf = m.nFaces();
    o.write(reinterpret_cast<const char *>(&nc), 4);
    o.write(reinterpret_cast<const char *>(&nb), 4);
    o.write(reinterpret_cast<const char *>(&nf), 4);
    auto wr = [&](double v) { o.write(reinterpret_cast<const char *>(&v), 8); };
    for (int c = 0; c < nc; ++c) wr(f.U[c].x);
    for (int c = 0; c < nc; ++c) wr(f.U[c].y);
    for (const ScalarField *s : {&f.p, &f.k, &f.omega, &f.nuT, &f.F1, &f.F2, &f.Pk, &f.CDkw})
        for (int c = 0; c < nc; ++c) wr((*s)[c]);
    for (int b = 0; b < nb; ++b) wr(f.U.bnd(b).x);
    for (int b = 0; b < nb; ++b) wr(f.U.bnd(b).y);
    for (const ScalarField *s : {&f.p, &f.k, &f.omega})
        for (int b = 0; b < nb; ++b) wr(s->bnd(b));
    for (int fi = 0; fi < nf; ++fi) wr(fi < (int)flux.size() ? flux[fi] : 0.0);
}

} // namespace

int main(int argc, char **argv) {
    if (argc < 2) {
        std::cerr << "usage: ref_harness <case.txt>\n";
        return 2;
    }
    std::ifstream in(argv[1]);
    if (!in) {
        std::cerr << "cannot open " << argv[1] << "\n";
        return 2;
    }
    Mesh mesh;
    bool haveMesh = false;
    double nu = 1.0;
    SolverSettings s;
    s.verbose = false;
    FlowBoundaryConditions bcs;
    bool haveBcs = false;
    std::string bcKind = "bcs_channel";
    double bcU = 1, bcK = 1e-4, bcW = 1;
    Vec3 U0(1, 0, 0);
    double p0 = 0, k0 = 1e-4, w0 = 1;
    ObservationOperator obs;
    std::vector<int> params;
    std::vector<std::vector<double>> thetas;
    std::string dumpPrefix, savePath;
    long long warmBudget = 0;

    std::string line;
    while (std::getline(in, line)) {
        auto hash = line.find('#');
        if (hash != std::string::npos) line.erase(hash);
        std::istringstream ls(line);
        std::string cmd;
        if (!(ls >> cmd)) continue;
        if (cmd == "mesh_file") {
            std::string p;
            ls >> p;
            mesh = Mesh::loadFromFile(p);
            haveMesh = true;
        } else if (cmd == "mesh_periodic_channel") {
            int nx, ny;
            double Lx, h, re, yp;
            ls >> nx >> ny >> Lx >> h >> re >> yp;
            mesh = Mesh::makePeriodicChannel2D(nx, ny, Lx, h, re, yp);
            haveMesh = true;
        } else if (cmd == "mesh_curved") {
            int n;
            ls >> n;
            std::vector<double> xs(n), ys(n);
            for (int i = 0; i < n; ++i) ls >> xs[i] >> ys[i];
            double yTop, re, yp;
            int ny;
            ls >> yTop >> ny >> re >> yp;
            mesh = Mesh::makeCurvedChannelPeriodic2D(xs, ys, yTop, ny, re, yp);
            haveMesh = true;
        } else if (cmd == "mesh_channel") {
            int nx, ny;
            double Lx, Ly;
            ls >> nx >> ny >> Lx >> Ly;
            mesh = Mesh::makeChannel2D(nx, ny, Lx, Ly);
            haveMesh = true;
        } else if (cmd == "mesh_channel_graded") {
            int nx, ny;
            double Lx, Ly, Re, yplus;
            ls >> nx >> ny >> Lx >> Ly >> Re >> yplus;
            mesh = Mesh::makeChannel2D(nx, ny, Lx, Ly, Re, yplus);
            haveMesh = true;
        } else if (cmd == "mesh_plate") {
            int nx, ny;
            double Lx, Ly, Re, yplus;
            ls >> nx >> ny >> Lx >> Ly >> Re >> yplus;
            mesh = Mesh::makePlate2D(nx, ny, Lx, Ly, Re, yplus);
            haveMesh = true;
        } else if (cmd == "mesh_bfs") {
            int nxu, nxd, nyu, nyd;
            double Lu, Ld, h, H, Re, yplus;
            ls >> nxu >> nxd >> nyu >> nyd >> Lu >> Ld >> h >> H >> Re >> yplus;
            mesh = Mesh::makeBackwardFacingStep2D(nxu, nxd, nyu, nyd, Lu, Ld, h, H, Re, yplus);
            haveMesh = true;
        } else if (cmd == "patch_type") {
            std::string name, type;
            ls >> name >> type;
            mesh.setPatchType(name, type);
        } else if (cmd == "save_mesh") {
            ls >> savePath;
        } else if (cmd == "nu") {
            ls >> nu;
        } else if (cmd == "bcs_channel") {
            ls >> bcU >> bcK >> bcW;
            haveBcs = true;
        } else if (cmd == "bcs_couette" || cmd == "bcs_plate") {
            ls >> bcU >> bcK >> bcW;
            haveBcs = true;
            bcKind = cmd;
        } else if (cmd == "set") {
            std::string k, v;
            ls >> k >> v;
            setKey(s, k, v);
        } else if (cmd == "init") {
            ls >> U0.x >> U0.y >> U0.z >> p0 >> k0 >> w0;
        } else if (cmd == "obs_u") {
            double x, y, o, sg;
            int c;
            ls >> x >> y >> c >> o >> sg;
            obs.addVelocityProfile(Vec3(x, y, 0.0), c, o, sg);
        } else if (cmd == "obs_cf") {
            std::string patch;
            double x, y, o, sg, rv;
            ls >> patch >> x >> y >> o >> sg >> rv;
            obs.addSkinFriction(patch, Vec3(x, y, 0.0), o, sg, rv);
        } else if (cmd == "obs_p") {
            double x, y, o, sg, rv, rp, rr;
            ls >> x >> y >> o >> sg >> rv >> rp >> rr;
            obs.addPressureTap(Vec3(x, y, 0.0), o, sg, rv, 0.0, rp, rr);
        } else if (cmd == "obs_drag") {
            std::string patch;
            double o, sg, ra, rv;
            ls >> patch >> o >> sg >> ra >> rv;
            obs.addDrag(patch, o, sg, ra, rv);
        } else if (cmd == "obs_sep" || cmd == "obs_reatt") {
            std::string patch;
            double o, sg;
            ls >> patch >> o >> sg;
            if (cmd == "obs_sep") obs.addSeparationPoint(patch, o, sg);
            else obs.addReattachmentLength(patch, o, sg);
        } else if (cmd == "params") {
            int v;
            while (ls >> v) params.push_back(v);
        } else if (cmd == "theta") {
            std::vector<double> row;
            double v;
            while (ls >> v) row.push_back(v);
            thetas.push_back(row);
        } else if (cmd == "dump") {
            ls >> dumpPrefix;
        } else if (cmd == "warm_cache") {
            ls >> warmBudget;
        } else {
            std::cerr << "unknown directive " << cmd << "\n";
            return 2;
        }
    }
    if (!haveMesh) {
        std::cerr << "no mesh\n";
        return 2;
    }
    if (!savePath.empty()) mesh.saveBinary(savePath);
    mesh.computeWallDistance();
    if (!haveBcs) bcU = 1.0;
    if (bcKind == "bcs_couette")
        bcs = FlowBoundaryConditions::couetteDefaults(mesh, bcU, bcK, bcW);
    else if (bcKind == "bcs_plate")
        bcs = FlowBoundaryConditions::flatPlateDefaults(mesh, bcU, bcK, bcW);
    else
        bcs = FlowBoundaryConditions::channelDefaults(mesh, bcU, bcK, bcW);
    InferenceParameterSet ps = InferenceParameterSet::fromIndices("harness", params.empty() ? std::vector<int>{9, 8} : params);

    ForwardModel model(mesh, ps, obs, bcs, nu, s, U0, p0, k0, w0);
    model.cache().setByteBudget((size_t)warmBudget);
    std::cout << "{\"n_cells\": " << mesh.nCells() << ", \"n_faces\": " << mesh.nFaces()
              << ", \"n_internal\": " << mesh.nInternalFaces() << ", \"rows\": [\n";
    for (size_t r = 0; r < thetas.size(); ++r) {
        auto t0 = std::chrono::steady_clock::now();
        ForwardEvaluationResult ev = model.evaluateWithDiagnostics(thetas[r], !dumpPrefix.empty());
        double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        const auto &res = ev.result;
        std::cout << "{\"status\": " << static_cast<int>(res.status) << ", \"status_name\": \""
                  << statusString(res.status) << "\", \"iters\": " << res.simpleIters
                  << ", \"loglik\": " << num(res.loglik)
                  << ", \"final_worst_measure\": " << num(ev.convergence.finalWorstMeasure)
                  << ", \"contraction_rate_u\": " << num(ev.convergence.contractionRateU)
                  << ", \"stalled\": " << (ev.convergence.stalled ? "true" : "false")
                  << ", \"converged\": " << (ev.convergence.converged ? "true" : "false")
                  << ", \"diverged\": " << (ev.convergence.diverged ? "true" : "false")
                  << ", \"used_warm_start\": " << (ev.usedWarmStart ? "true" : "false")
                  << ", \"wall_seconds\": " << num(wall) << ", \"predictions\": [";
        for (size_t i = 0; i < res.predictions.size(); ++i)
            std::cout << (i ? ", " : "") << num(res.predictions[i]);
        std::cout << "]";
        if (model.hasLastFields()) {
            const FlowFields &f = model.lastFields();
            double ux = 0, uy = 0, p = 0, k = 0, w = 0, nut = 0;
            for (int ci = 0; ci < mesh.nCells(); ++ci) {
                ux += f.U[ci].x * f.U[ci].x;
                uy += f.U[ci].y * f.U[ci].y;
                p += f.p[ci] * f.p[ci];
                k += f.k[ci] * f.k[ci];
                w += f.omega[ci] * f.omega[ci];
                nut += f.nuT[ci] * f.nuT[ci];
            }
            std::cout << ", \"field_norms\": {\"sumsq_ux\": " << num(ux) << ", \"sumsq_uy\": "
                      << num(uy) << ", \"sumsq_p\": " << num(p) << ", \"sumsq_k\": " << num(k)
                      << ", \"sumsq_omega\": " << num(w) << ", \"sumsq_nut\": " << num(nut) << "}";
            if (!dumpPrefix.empty() && ev.snapshot)
                dumpFields(dumpPrefix + "_" + std::to_string(r) + ".bin", mesh, f,
                           ev.snapshot->faceFluxes);
        }
        std::cout << "}" << (r + 1 < thetas.size() ? "," : "") << "\n";
    }
    std::cout << "]}\n";
    return 0;
}
