// VectorSolver driver for the coupled Allen-Cahn form (config 3 family):
//   ac_solver <in.vfield> <outdir> <nx> <ny> <nsteps> <bcmode>
// component i: (L*kappa)*lap(eta_i) + pw(eta_i, e - e^3, L)
//              + sum_{j != i} pw(eta_i, eta_j, a*b*b, -2*L*gamma)
// Writes ac_<tag>.vfield for the fused plan, the generic plan (an extra zero
// term), the host path and FP32 bulk.
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <string>
#include <vector>

#include "boundary/NoFluxBC.h"
#include "boundary/PeriodicBC.h"
#include "equation/VectorEquation.h"
#include "mesh/Mesh.h"
#include "solver/VectorSolver.h"

using namespace PhiX;

int main(int argc, char** argv) {
    if (argc < 7) {
        std::fprintf(stderr, "usage: ac_solver in.vfield outdir nx ny nsteps bcmode\n");
        return 1;
    }
    try {
        const std::string in = argv[1], out = argv[2];
        const int nx = std::atoi(argv[3]), ny = std::atoi(argv[4]), nsteps = std::atoi(argv[5]);
        const std::string bcmode = argv[6];
        Mesh mesh = Mesh::makeUniform2D(CoordSys::CARTESIAN, nx, 1.0, 0.0, ny, 1.0, 0.0);
        const double L = 1.0, kappa = 2.0, gamma = 1.5, dt = 0.01;
        std::unique_ptr<BoundaryCondition> b0, b1;
        if (bcmode == "periodic") {
            b0.reset(new PeriodicBC(Axis::X));
            b1.reset(new PeriodicBC(Axis::Y));
        } else {
            b0.reset(new NoFluxBC(Axis::X));
            b1.reset(new NoFluxBC(Axis::Y));
        }
        std::vector<BoundaryCondition*> bcs{b0.get(), b1.get()};
        for (int mode = 0; mode < 4; ++mode) {
            const char* tag = mode == 0 ? "fused" : (mode == 1 ? "generic" : (mode == 2 ? "cpu" : "fp32"));
            VectorField in_vf = VectorField::readFromFile(mesh, in, 1);
            const int N = in_vf.nComponents();
            VectorField eta(mesh, "eta", N, 1);
            for (int c = 0; c < N; ++c) {
                eta[c].curr = in_vf[c].curr;
                eta[c].prev = in_vf[c].curr;
            }
            eta.allocDevice();
            eta.uploadAllToDevice();
            VectorEquation veq(eta, "AC");
            VectorRHSExpr rhs(N);
            for (int i = 0; i < N; ++i) {
                RHSExpr e((L * kappa) * lap(eta[i]));
                e += pw(eta[i], PHIX_FN(double v) { return v - v * v * v; }, L);
                for (int j = 0; j < N; ++j)
                    if (j != i) e += pw(eta[i], eta[j], PHIX_FN(double a, double b) { return a * b * b; }, -2.0 * L * gamma);
                if (mode == 1) e += pw(eta[i], PHIX_FN(double) { return 0.0; });
                rhs[i] = e;
            }
            veq.setRHS(rhs);
            VectorSolver vs(veq, bcs, dt, TimeScheme::RK4);
            if (mode == 3) vs.setBulkPrecision(Precision::FP32);
            std::printf("plan %s: %s\n", tag, vs.plan().c_str());
            if (mode == 2) {
                for (int s = 0; s < nsteps; ++s) vs.advanceCPU();
            } else {
                vs.run(nsteps);
                eta.downloadAllFromDevice();
            }
            eta.write(out + "/ac_" + tag + ".vfield");
        }
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 2;
    }
    return 0;
}
