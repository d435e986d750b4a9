// CH_AC_2D -- coupled Cahn-Hilliard / Allen-Cahn (Ostwald ripening) driven by
// a JSONC settings file; the counterpart of the reference's CH_AC_2D program
// (applications/solvers/Cahn-Hillard+Allen-Cahn_double-well/2D/CH+AC_double-well.cu).
//
//   ch_ac_2d [path/to/settings.jsonc]       (default settings/settings.jsonc)
//
//   mu      = 2 rho^2 (c - ca) + 2 rho^2 (ca - cb) h(eta) - kappa_c lap(c)      (STEADY)
//   h(eta)  = eta^3 (6 eta^2 - 15 eta + 10)
//   dc/dt   = M lap(mu)                                                      (TRANSIENT)
//   deta/dt = -L [30 rho^2 eta^2 (1 - eta)^2 (2c - ca - cb)(ca - cb)
//                 + 2 w eta (1 - eta)(1 - 2 eta)] + L kappa_eta lap(eta)      (TRANSIENT)
//
// Settings blocks as for ch_2d, with constants {M, rho, ca, cb, kappa_c,
// kappa_eta, L, w}; initial fields c, eta and mu. The equations run on the
// generic device path (one kernel per term), the mu update before both
// transient equations of a step, as the reference's multi-step Solver orders
// them.
#include <iostream>
#include <string>

#include "IO/ConfigFile.h"
#include "IO/FieldIO.h"
#include "IO/OutputWriter.h"
#include "boundary/BCFactory.h"
#include "equation/Equation.h"
#include "settings_run.h"

int main(int argc, char* argv[]) {
    using namespace PhiX;
    try {
        const IO::ConfigFile cfg = IO::ConfigFile::fromArgs(argc, argv);
        const IO::Json& init = cfg["initialize"];
        const IO::Json& k = cfg["constants"];

        Mesh mesh = apps::meshFrom(cfg["mesh"]);
        mesh.print();
        const double dt = init["dt"];
        const int nSteps = init["nSteps"];
        const double rho = k["rho"], ca = k["ca"], cb = k["cb"], M = k["M"], L = k["L"], w = k["w"];
        const double kappa_c = k["kappa_c"], kappa_eta = k["kappa_eta"];

        ScalarField c(mesh, "c", 1), mu(mesh, "mu", 1), eta(mesh, "eta", 1);
        for (ScalarField* f : {&c, &mu, &eta}) f->fill(0);
        const int start = IO::resolveStartStep(init["start_from"].get<std::string>());
        IO::initField(c, start);
        IO::initField(eta, start);
        if (start == 0) IO::initField(mu, start);
        for (ScalarField* f : {&c, &eta, &mu}) {
            f->allocDevice();
            f->uploadAllToDevice();
        }

        BCSet bc = buildBCs(cfg["boundary_conditions"]);

        Equation eqMu(mu, "CH_mu");
        eqMu.setRHS(pw(c, PHIX_FN(double v) { return 2.0 * rho * rho * (v - ca); }) +
                    pw(eta, PHIX_FN(double e) {
                        const double h = e * e * e * (6.0 * e * e - 15.0 * e + 10.0);
                        return 2.0 * rho * rho * (ca - cb) * h;
                    }) -
                    kappa_c * lap(c));
        Equation eqC(c, "CH_c");
        eqC.setRHS(M * lap(mu));
        Equation eqEta(eta, "AC_eta");
        eqEta.setRHS(pw(eta, c,
                        PHIX_FN(double e, double v) {
                            const double bulk =
                                30.0 * rho * rho * e * e * (1.0 - e) * (1.0 - e) * (2.0 * v - ca - cb) * (ca - cb);
                            const double well = 2.0 * w * e * (1.0 - e) * (1.0 - 2.0 * e);
                            return -L * (bulk + well);
                        }) +
                     L * kappa_eta * lap(eta));

        Solver solver({{&c, bc.ptrs, &eqMu, EquationType::STEADY},
                       {&mu, bc.ptrs, &eqC, EquationType::TRANSIENT},
                       {&eta, bc.ptrs, &eqEta, EquationType::TRANSIENT}},
                      dt, TimeScheme::EULER);
        solver.step = start;
        solver.time = start * dt;
        std::cout << "  solver plan: " << solver.plan() << "\n";

        IO::OutputWriter writer(cfg["output"]);
        if (start == 0) {
            writer.writeFields(c, 0, solver.time);
            writer.writeFields(eta, 0, solver.time);
            std::cout << "Starting CH+AC simulation (" << nSteps << " steps, dt=" << dt << ")\n";
        } else {
            std::cout << "Resuming CH+AC simulation from step " << start << " (t=" << start * dt << "), "
                      << nSteps - start << " steps remaining, dt=" << dt << "\n";
        }
        writer.resetTimer();
        apps::timeLoop(solver, writer, {&c, &eta}, nSteps);
        std::cout << "Done.\n";
        return 0;
    } catch (const std::exception& e) {
        std::cerr << "ch_ac_2d: " << e.what() << "\n";
        return 1;
    }
}
