// Prints the Fe-B model functions of phix_amd/apps/gfa_model.h, evaluated on
// the host, one "name args... value" line each, over temperatures that reach
// every SGTE piece (both sides of each breakpoint), both branches of f(tau),
// and B fractions inside and outside the clamp. tests/test_oracle_gfa.py
// compares every line with the oracle.
#include <cmath>
#include <cstdio>

#include "../../phix_amd/apps/gfa_model.h"

using namespace PhiX::gfa;

int main() {
    const double temps[] = {300.0, 450.0, 500.0, 501.0, 1000.0, 1100.0, 1101.0, 1811.0, 1812.0,
                            1900.0, 2348.0, 2349.0, 2500.0, 3000.0, 3001.0, 3100.0};
    for (double T : temps) {
        std::printf("G_B_liquid %.17g %.17g\n", T, gLiquidB(T));
        std::printf("G_Fe_liquid %.17g %.17g\n", T, gLiquidFe(T));
        std::printf("L_BFe_liquid %.17g %.17g\n", T, lLiquidBFe(T));
        std::printf("G_Fe_bcc %.17g %.17g\n", T, gBccFe(T));
        std::printf("G_B_beta %.17g %.17g\n", T, gBetaB(T));
        std::printf("deltaG_Fe3B %.17g %.17g\n", T, dgFormFe3B(T));
        std::printf("f_S %.17g %.17g %.17g\n", T, 2.44e-5, fSolid(T, 2.44e-5));
        std::printf("df_AmL %.17g %.17g %.17g %.17g %.17g\n", T, 710.0, 0.45, 2.64e-5,
                    dfAmorphous(T, 710.0, 0.45, 2.64e-5));
        for (double c : {-0.1, 0.0, 1e-13, 0.05, 0.25, 0.33, 0.5, 0.9, 1.0, 1.3}) {
            std::printf("dGL_dc %.17g %.17g %.17g\n", c, T, dGLiquidDc(c, T));
            std::printf("G_L %.17g %.17g %.17g\n", c, T, gLiquid(c, T));
        }
    }
    for (double tau : {0.25, 0.5, 0.9, 1.0, 1.0000001, 1.408, 2.0, 5.0}) std::printf("f_tau %.17g %.17g\n", tau, indenF(tau));
    for (double x : {-0.1, 0.0, 0.2, 0.5, 0.77, 1.0, 1.1}) {
        std::printf("h %.17g %.17g\n", x, hInterp(x));
        std::printf("h_prime %.17g %.17g\n", x, hInterpPrime(x));
        std::printf("g_prime %.17g %.17g\n", x, wellPrime(x));
    }
    return 0;
}
