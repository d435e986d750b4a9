// Drives the drop-in C++ API the way the reference's CH applications do and
// writes the resulting fields in the .field format for comparison with the
// oracle (tests/test_gpu_cpp_api.py).
//
//   ch_solver euler <in.field> <outdir> <nx> <ny> <nsteps>
//       2D periodic, multi-step {c, mu STEADY}, {mu, c TRANSIENT} Euler
//       (applications/solvers/Cahn-Hillard_double-well/2D/Cahn-Hillard_double-well.cu:84-135)
//   ch_solver rk4 <in.field> <outdir> <nx> <ny> <nz> <nsteps> <bcmode>
//       single-equation RK4, RHS = M*lap(pw(c,f') - kappa*lap(c), bcs)
//       bcmode: noflux | periodic | mixed
// Each run writes <tag>_c.field (and _mu) for the fused plan, the generic plan
// (forced by an extra zero term the planner does not match), the host path
// (advanceCPU) and, for rk4, the FP32-bulk fused plan. Plan names go to stdout.
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <string>
#include <vector>

#include "boundary/FixedBC.h"
#include "boundary/NoFluxBC.h"
#include "boundary/PeriodicBC.h"
#include "equation/Equation.h"
#include "mesh/Mesh.h"
#include "solver/Solver.h"

using namespace PhiX;

static void load(ScalarField& f, const Mesh& m, const std::string& path) {
    ScalarField in = ScalarField::readFromFile(m, path, f.ghost);
    f.curr = in.curr;
    f.prev = in.curr;
}

static int euler(int argc, char** argv) {
    const std::string in = argv[2], out = argv[3];
    const int nx = std::atoi(argv[4]), ny = std::atoi(argv[5]), nsteps = std::atoi(argv[6]);
    Mesh mesh = Mesh::makeUniform2D(CoordSys::CARTESIAN, nx, 1.0, 0.0, ny, 1.0, 0.0);
    const double rho = 1.0, ca = 0.0, cb = 1.0, kappa = 0.5, M = 1.0, dt = 0.01;
    PeriodicBC bx(Axis::X), by(Axis::Y);
    std::vector<BoundaryCondition*> bcs{&bx, &by};
    for (int mode = 0; mode < 3; ++mode) {
        const char* tag = mode == 0 ? "fused" : (mode == 1 ? "generic" : "cpu");
        ScalarField c(mesh, "c", 1), mu(mesh, "mu", 1);
        load(c, mesh, in);
        c.allocDevice();
        c.uploadAllToDevice();
        mu.allocDevice();
        mu.uploadAllToDevice();
        Equation eqMu(mu, "CH_mu");
        RHSExpr e = pw(c, PHIX_FN(double v) { return 2.0 * rho * (v - ca) * (v - cb) * (2.0 * v - ca - cb); }) -
                    kappa * lap(c);
        if (mode == 1) e += pw(c, PHIX_FN(double) { return 0.0; });
        eqMu.setRHS(e);
        Equation eqC(c, "CH_c");
        eqC.setRHS(M * lap(mu));
        Solver solver({{&c, bcs, &eqMu, EquationType::STEADY}, {&mu, bcs, &eqC, EquationType::TRANSIENT}}, dt,
                      TimeScheme::EULER);
        std::printf("plan %s: %s\n", tag, solver.plan().c_str());
        if (mode == 2) {
            for (int s = 0; s < nsteps; ++s) solver.advanceCPU();
        } else {
            for (int s = 0; s < nsteps; ++s) solver.advance();
            c.downloadAllFromDevice();
            mu.downloadCurrFromDevice();
            bool prev_ok = c.prev == c.curr;
            std::printf("prev==curr %s: %d\n", tag, (int)prev_ok);
        }
        std::printf("step %s: %d time %.17g\n", tag, solver.step, solver.time);
        c.write(out + "/euler_" + tag + "_c.field");
        mu.write(out + "/euler_" + tag + "_mu.field");
    }
    return 0;
}

static int rk4(int argc, char** argv) {
    const std::string in = argv[2], out = argv[3];
    const int nx = std::atoi(argv[4]), ny = std::atoi(argv[5]), nz = std::atoi(argv[6]), nsteps = std::atoi(argv[7]);
    const std::string bcmode = argv[8];
    Mesh mesh = nz > 1 ? Mesh::makeUniform3D(CoordSys::CARTESIAN, nx, 1.0, 0.0, ny, 1.0, 0.0, nz, 1.0, 0.0)
                       : Mesh::makeUniform2D(CoordSys::CARTESIAN, nx, 1.0, 0.0, ny, 1.0, 0.0);
    const double rho = 1.0, ca = 0.0, cb = 1.0, kappa = 1.0, M = 1.0, dt = 0.005;
    std::vector<std::unique_ptr<BoundaryCondition>> own;
    for (int a = 0; a < mesh.dim; ++a) {
        Axis ax = (Axis)a;
        if (bcmode == "periodic" || (bcmode == "mixed" && a == 1)) own.emplace_back(new PeriodicBC(ax));
        else if (bcmode == "mixed" && a == 0) {
            own.emplace_back(new FixedBC(ax, Side::LOW, 0.45));
            own.emplace_back(new NoFluxBC(ax, Side::HIGH));
        } else own.emplace_back(new NoFluxBC(ax));
    }
    std::vector<BoundaryCondition*> bcs;
    for (auto& b : own) bcs.push_back(b.get());
    for (int mode = 0; mode < 4; ++mode) {
        const char* tag = mode == 0 ? "fused" : (mode == 1 ? "generic" : (mode == 2 ? "cpu" : "fp32"));
        ScalarField c(mesh, "c", 1);
        load(c, mesh, in);
        c.allocDevice();
        c.uploadAllToDevice();
        Equation eq(c, "CH");
        RHSExpr mu = pw(c, PHIX_FN(double v) { return 2.0 * rho * (v - ca) * (v - cb) * (2.0 * v - ca - cb); }) -
                     kappa * lap(c);
        if (mode == 1) mu += pw(c, PHIX_FN(double) { return 0.0; });
        eq.setRHS(M * lap(mu, bcs));
        Solver solver(eq, bcs, dt, TimeScheme::RK4);
        if (mode == 3) solver.setBulkPrecision(Precision::FP32);
        std::printf("plan %s: %s\n", tag, solver.plan().c_str());
        if (mode == 2) {
            for (int s = 0; s < nsteps; ++s) solver.advanceCPU();
        } else {
            int calls = 0;
            solver.run(nsteps, 7, [&](const Solver&) { ++calls; });
            c.downloadAllFromDevice();
            std::printf("prev==curr %s: %d callbacks %d\n", tag, (int)(c.prev == c.curr), calls);
        }
        c.write(out + "/rk4_" + tag + "_c.field");
    }
    return 0;
}

int main(int argc, char** argv) {
    try {
        if (argc >= 7 && std::string(argv[1]) == "euler") return euler(argc, argv);
        if (argc >= 9 && std::string(argv[1]) == "rk4") return rk4(argc, argv);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 2;
    }
    std::fprintf(stderr, "usage: ch_solver euler|rk4 ...\n");
    return 1;
}
