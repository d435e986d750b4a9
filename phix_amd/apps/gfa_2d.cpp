// GFA_2D -- glass-formation ability of Fe-B (crystal order phi, amorphous
// order eta, B fraction c) driven by a JSONC settings file; the counterpart of
// the reference's GFA_2D program
// (applications/solvers/glass_formation/2D/GFA.cu:87-233).
//
//   gfa_2d [path/to/settings.jsonc]       (default settings/settings.jsonc)
//
// One Euler step is six multi-steps, in the reference's order (GFA.cu:192-201):
//   1. BCs(c);   mu  = (1 - h(phi)) dG_L/dc(c, T) / Vm_L                  STEADY
//   2. BCs(mu);  D   = M_c(phi, eta) c (1 - c)                             STEADY
//   3. BCs(D);   D   recomputed (which clears D's ghosts again: step 4    STEADY
//                reads D = 0 beyond the faces, as the reference does)
//   4. BCs(c);   c  += dt [D lap(mu) + dD/dx dmu/dx + dD/dy dmu/dy]        TRANSIENT
//   5. BCs(phi); phi += dt [-M_phi dF/dphi + M_phi eps^2 lap(phi)]         TRANSIENT
//   6. BCs(eta); eta += dt [-M_eta dF/deta + M_eta beta^2 lap(eta)]        TRANSIENT
// with M_c = [(1 - h(phi))((1 - h(eta)) D_L + h(eta) D_Am) + h(phi) D_S] / (R T),
// dF/dphi = h'(phi)(f_S - f_L(c) - h(eta) df_Am) + w_phi g'(phi) + 2 w_ex phi eta^2,
// dF/deta = (1 - h(phi)) h'(eta) df_Am + w_eta g'(eta) + 2 w_ex phi^2 eta,
// and the Arrhenius mobilities M_phi, D_L, D_S of the temperature T.
//
// Settings: "mesh" {nx, ny, dx, dy, x0, y0}; "initialize" {start_from, dt,
// nSteps}; "boundary_conditions"; "constants" {M_eta, eps_sq, beta_sq, w_phi,
// w_eta, w_ex, T, T_g, alpha, V_m_L, V_m_S, D_Am}; "output". Initial fields
// c, phi, eta and (cold start) mu.
//
// By default a step is one fused launch (phx_gfa_euler_steps: mu and D on
// chip, c, phi, eta read and written once). "initialize": {"solver":
// "generic"} (absent in the reference's settings) runs the six equations as
// written below on the generic per-term device path instead; both give the
// same fields to rounding.
#include <algorithm>
#include <iostream>
#include <string>

#include <hip/hip_runtime.h>

#include "IO/ConfigFile.h"
#include "IO/FieldIO.h"
#include "IO/OutputWriter.h"
#include "boundary/BCFactory.h"
#include "equation/Equation.h"
#include "gfa_model.h"
#include "phix_amd.h"
#include "settings_run.h"

namespace {

// The fused path's grid: the mesh, the fields' ghost width, and the faces of
// the settings' "boundary_conditions" (buildBCs has accepted them: Periodic
// or NoFlux, periodic on both sides of an axis or neither).
phx_grid fusedGrid(const PhiX::Mesh& mesh, int ghost, const PhiX::IO::Json& b) {
    phx_grid g{};
    g.dim = mesh.dim;
    for (int a = 0; a < 3; ++a) {
        g.n[a] = mesh.n[a];
        g.d[a] = mesh.d[a];
    }
    g.ghost = ghost;
    const char* faces[2][2] = {{"x_min", "x_max"}, {"y_min", "y_max"}};
    for (int a = 0; a < 2; ++a)
        for (int s = 0; s < 2; ++s)
            g.bc[a][s] = b[faces[a][s]].get<std::string>() == "Periodic" ? PHX_BC_PERIODIC : PHX_BC_NOFLUX;
    return g;
}

void hipCheck(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

}  // namespace

int main(int argc, char* argv[]) {
    using namespace PhiX;
    using namespace PhiX::gfa;
    try {
        const IO::ConfigFile cfg = IO::ConfigFile::fromArgs(argc, argv);
        const IO::Json& init = cfg["initialize"];
        const IO::Json& k = cfg["constants"];

        Mesh mesh = apps::meshFrom(cfg["mesh"]);
        mesh.print();
        const double dt = init["dt"];
        const int nSteps = init["nSteps"];

        const double M_eta = k["M_eta"], eps_sq = k["eps_sq"], beta_sq = k["beta_sq"];
        const double w_phi = k["w_phi"], w_eta = k["w_eta"], w_ex = k["w_ex"];
        const double T = k["T"], T_g = k["T_g"], alpha = k["alpha"];
        const double V_m_L = k["V_m_L"], V_m_S = k["V_m_S"], D_Am = k["D_Am"];

        // Temperature-only quantities, fixed for the run (GFA.cu:117-121).
        const double M_phi = 22.1 * exp(-140.0e3 / (kR * T));
        const double D_L = 2.0e-6 * exp(-1.11e5 / (kR * T));
        const double D_S = 1.311e-6 * exp(-1.51e5 / (kR * T));
        const double fS = fSolid(T, V_m_S);
        const double dfAm = dfAmorphous(T, T_g, alpha, V_m_L);

        ScalarField c(mesh, "c", 1), phi(mesh, "phi", 1), eta(mesh, "eta", 1), mu(mesh, "mu", 1), D(mesh, "D", 1);
        for (ScalarField* f : {&c, &phi, &eta, &mu, &D}) f->fill(0);
        const int start = IO::resolveStartStep(init["start_from"].get<std::string>());
        IO::initField(c, start);
        IO::initField(phi, start);
        IO::initField(eta, start);
        if (start == 0) IO::initField(mu, start);
        for (ScalarField* f : {&c, &phi, &eta, &mu, &D}) {
            f->allocDevice();
            f->uploadAllToDevice();
        }

        BCSet bc = buildBCs(cfg["boundary_conditions"]);

        Equation eqMu(mu, "mu");
        eqMu.setRHS(pw(c, phi, PHIX_FN(double cv, double pv) {
            return (1.0 - hInterp(pv)) * dGLiquidDc(cv, T) / V_m_L;
        }));

        Equation eqD(D, "D");
        eqD.setRHS(pw(c, phi, eta, PHIX_FN(double cv, double pv, double ev) {
            const double hp = hInterp(pv), he = hInterp(ev);
            const double Mc = ((1.0 - hp) * ((1.0 - he) * D_L + he * D_Am) + hp * D_S) / (kR * T);
            return Mc * cv * (1.0 - cv);
        }));

        Equation eqC(c, "CH_c");
        eqC.setRHS(D * lap(mu) + grad(D, 0) * grad(mu, 0) + grad(D, 1) * grad(mu, 1));

        Equation eqPhi(phi, "AC_phi");
        eqPhi.setRHS(pw(phi, eta, c,
                        PHIX_FN(double pv, double ev, double cv) {
                            const double dF = hInterpPrime(pv) * (fS - fLiquid(cv, T, V_m_L) - hInterp(ev) * dfAm) +
                                              w_phi * wellPrime(pv) + 2.0 * w_ex * pv * ev * ev;
                            return -M_phi * dF;
                        }) +
                     M_phi * eps_sq * lap(phi));

        Equation eqEta(eta, "AC_eta");
        eqEta.setRHS(pw(phi, eta,
                        PHIX_FN(double pv, double ev) {
                            const double dF = (1.0 - hInterp(pv)) * hInterpPrime(ev) * dfAm +
                                              w_eta * wellPrime(ev) + 2.0 * w_ex * pv * pv * ev;
                            return -M_eta * dF;
                        }) +
                     M_eta * beta_sq * lap(eta));

        const std::string plan = init.value("solver", "fused");
        if (plan != "fused" && plan != "generic")
            throw std::runtime_error("initialize.solver must be \"fused\" or \"generic\", got \"" + plan + "\"");
        if (plan == "fused") {
            const phx_grid g = fusedGrid(mesh, c.ghost, cfg["boundary_conditions"]);
            const phx_gfa_params p{dt, M_eta, eps_sq, beta_sq, w_phi, w_eta, w_ex, T, T_g, alpha, V_m_L, V_m_S, D_Am};
            void* work = nullptr;
            hipCheck(hipMalloc(&work, phx_gfa_workspace_bytes(&g)), "gfa_2d: workspace");
            std::cout << "  solver plan: fused GFA step (phx_gfa_euler_steps)\n";
            IO::OutputWriter writer(cfg["output"]);
            int step = start;
            double time = start * dt;
            if (start == 0) {
                for (ScalarField* f : {&c, &phi, &eta}) writer.writeFields(*f, 0, time);
                std::cout << "Starting GFA simulation (" << nSteps << " steps, dt=" << dt << ", T=" << T << " K)\n";
            } else {
                std::cout << "Resuming GFA simulation from step " << start << " (t=" << start * dt << "), "
                          << nSteps - start << " steps remaining, dt=" << dt << "\n";
            }
            writer.resetTimer();
            // the same cadence as apps::timeLoop: one device call up to the next
            // print or write step; time accumulates dt per step as Solver does
            auto nextMultiple = [&](int every) { return (step / every + 1) * every; };
            while (step < nSteps) {
                const int stop = std::min({nextMultiple(writer.printInterval), nextMultiple(writer.writeInterval), nSteps});
                if (phx_gfa_euler_steps(&g, c.d_curr, phi.d_curr, eta.d_curr, work, &p, stop - step, nullptr) != PHX_OK)
                    throw std::runtime_error(std::string("phx_gfa_euler_steps: ") + phx_last_error());
                hipCheck(hipDeviceSynchronize(), "gfa_2d");
                for (; step < stop; ++step) time += dt;
                if (writer.shouldPrint(step)) writer.printProgress(step, time);
                if (writer.shouldWrite(step))
                    for (ScalarField* f : {&c, &phi, &eta}) writer.writeFields(*f, step, time);
            }
            hipCheck(hipFree(work), "gfa_2d: workspace");
            std::cout << "Done.\n";
            return 0;
        }

        Solver solver({{&c, bc.ptrs, &eqMu, EquationType::STEADY},
                       {&mu, bc.ptrs, &eqD, EquationType::STEADY},
                       {&D, bc.ptrs, &eqD, EquationType::STEADY},
                       {&c, bc.ptrs, &eqC, EquationType::TRANSIENT},
                       {&phi, bc.ptrs, &eqPhi, EquationType::TRANSIENT},
                       {&eta, bc.ptrs, &eqEta, EquationType::TRANSIENT}},
                      dt, TimeScheme::EULER);
        solver.step = start;
        solver.time = start * dt;
        std::cout << "  solver plan: " << solver.plan() << "\n";

        IO::OutputWriter writer(cfg["output"]);
        if (start == 0) {
            for (ScalarField* f : {&c, &phi, &eta}) writer.writeFields(*f, 0, solver.time);
            std::cout << "Starting GFA simulation (" << nSteps << " steps, dt=" << dt << ", T=" << T << " K)\n";
        } else {
            std::cout << "Resuming GFA simulation from step " << start << " (t=" << start * dt << "), "
                      << nSteps - start << " steps remaining, dt=" << dt << "\n";
        }
        writer.resetTimer();
        apps::timeLoop(solver, writer, {&c, &phi, &eta}, nSteps);
        std::cout << "Done.\n";
        return 0;
    } catch (const std::exception& e) {
        std::cerr << "gfa_2d: " << e.what() << "\n";
        return 1;
    }
}
