// CH_2D -- Cahn-Hilliard with the double-well bulk energy, driven by a JSONC
// settings file; the counterpart of the reference's CH_2D program
// (applications/solvers/Cahn-Hillard_double-well/2D/Cahn-Hillard_double-well.cu).
//
//   ch_2d [path/to/settings.jsonc]          (default settings/settings.jsonc)
//
//   mu    = 2 rho (c - ca)(c - cb)(2c - ca - cb) - kappa lap(c)   (STEADY)
//   dc/dt = M lap(mu)                                             (TRANSIENT)
//
// Settings blocks: mesh {nx, ny, dx, dy, x0, y0}, initialize {start_from, dt,
// nSteps}, boundary_conditions {x_min ... y_max}, constants {M, kappa, rho, ca,
// cb}, output {print_interval, write_interval, format}. A cold start reads
// settings/initial_field/{c,mu}.field, a warm start output/c_<step>.field, all
// relative to the working directory. With periodic or no-flux BCs the solver
// plans the fused Euler kernels (phx_ch_euler_steps) for this equation pair.
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
        const double M = k["M"], kappa = k["kappa"], rho = k["rho"], ca = k["ca"], cb = k["cb"];

        ScalarField c(mesh, "c", 1), mu(mesh, "mu", 1);
        c.fill(0);
        mu.fill(0);
        const int start = IO::resolveStartStep(init["start_from"].get<std::string>());
        IO::initField(c, start);
        if (start == 0) IO::initField(mu, start);
        for (ScalarField* f : {&c, &mu}) {
            f->allocDevice();
            f->uploadAllToDevice();
        }

        BCSet bc = buildBCs(cfg["boundary_conditions"]);

        Equation eqMu(mu, "CH_mu");
        eqMu.setRHS(pw(c, PHIX_FN(double v) { return 2.0 * rho * (v - ca) * (v - cb) * (2.0 * v - ca - cb); }) -
                    kappa * lap(c));
        Equation eqC(c, "CH_c");
        eqC.setRHS(M * lap(mu));

        Solver solver({{&c, bc.ptrs, &eqMu, EquationType::STEADY}, {&mu, bc.ptrs, &eqC, EquationType::TRANSIENT}},
                      dt, TimeScheme::EULER);
        apps::applyPrecision(solver, init);
        solver.step = start;
        solver.time = start * dt;
        std::cout << "  solver plan: " << solver.plan() << "\n";

        IO::OutputWriter writer(cfg["output"]);
        if (start == 0) {
            writer.writeFields(c, 0, solver.time);
            std::cout << "Starting Cahn-Hilliard simulation (" << nSteps << " steps, dt=" << dt << ")\n";
        } else {
            std::cout << "Resuming Cahn-Hilliard simulation from step " << start << " (t=" << start * dt << "), "
                      << nSteps - start << " steps remaining, dt=" << dt << "\n";
        }
        writer.resetTimer();
        apps::timeLoop(solver, writer, {&c}, nSteps);
        std::cout << "Done.\n";
        return 0;
    } catch (const std::exception& e) {
        std::cerr << "ch_2d: " << e.what() << "\n";
        return 1;
    }
}
