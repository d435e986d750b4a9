#pragma once
// Fe-B glass-formation model of the reference's GFA_2D program
// (applications/solvers/glass_formation/2D/GFA.cu:19-85): SGTE-form Gibbs
// energies of the liquid, bcc-Fe, beta-B and Fe3B phases, the liquid mixing
// free energy, and the Inden-form glass-transition function. Callable on the
// host and on the device (the solver's pw() functors use them).
//
// Each SGTE piece is G(T) = a + b T + c T ln T + d T^2 + e T^3 + f / T
// + g7 T^7 + gm9 T^-9 on T <= tmax. The coefficients are the reference's,
// including the Fe 1/T coefficient 773597 (GFA.cu:32,39). SGTE's published
// GHSER_Fe has 77359 there; with 773597 the Fe pieces jump by 384.4 J/mol at
// 1811 K, while every other piece is continuous at its breakpoints. The
// reference's value is kept so results match its program.
#include <cmath>

namespace PhiX {
namespace gfa {

constexpr double kR = 8.314;   // J / (mol K)

struct SgtePiece {
    double a, b, c, d, e, f, g7, gm9;
};

__host__ __device__ inline double sgte(const SgtePiece& p, double T) {
    const double T2 = T * T, T3 = T2 * T;
    double g = p.a + p.b * T + p.c * T * log(T) + p.d * T2 + p.e * T3 + p.f / T;
    if (p.g7 != 0.0) g += p.g7 * (T3 * T3 * T);
    if (p.gm9 != 0.0) g += p.gm9 / (T3 * T3 * T3);
    return g;
}

// Liquid boron (GFA.cu:21-27).
__host__ __device__ inline double gLiquidB(double T) {
    if (T <= 500.0) return sgte({40723.275, 86.843839, -15.6641, -6.864515e-3, 0.618878e-6, 370843.0, 0, 0}, T);
    if (T <= 2348.0) return sgte({41119.703, 82.101722, -14.9827763, -7.095669e-3, 0.507347e-6, 335484.0, 0, 0}, T);
    return sgte({28842.012, 200.94731, -31.4, 0, 0, 0, 0, 0}, T);
}

// Liquid iron (GFA.cu:28-35).
__host__ __device__ inline double gLiquidFe(double T) {
    if (T <= 1811.0)
        return sgte({13265.87, 117.57557, -23.5143, -4.39752e-3, -0.058927e-6, 773597.0, -367.516e-23, 0}, T);
    return sgte({-10838.83, 291.302, -46.0, 0, 0, 0, 0, 0}, T);
}

// B-Fe liquid interaction parameter, zeroth order (GFA.cu:36).
__host__ __device__ inline double lLiquidBFe(double T) { return -122861.0 + 14.59 * T; }

// bcc iron (GFA.cu:37-42).
__host__ __device__ inline double gBccFe(double T) {
    if (T <= 1811.0) return sgte({1225.7, 124.134, -23.5143, -4.39752e-3, -0.058927e-6, 773597.0, 0, 0}, T);
    return sgte({-25383.581, 299.31255, -46.0, 0, 0, 0, 0, 2296.03e28}, T);
}

// beta boron (GFA.cu:43-51).
__host__ __device__ inline double gBetaB(double T) {
    if (T <= 1100.0) return sgte({-7735.284, 107.111864, -15.6641, -6.864515e-3, 0.618878e-6, 370843.0, 0, 0}, T);
    if (T <= 2348.0)
        return sgte({-16649.474, 184.801744, -26.6047, -0.79809e-3, -0.02556e-6, 1748270.0, 0, 0}, T);
    if (T <= 3000.0)
        return sgte({-36667.582, 231.336244, -31.5957527, -1.59488e-3, 0.134719e-6, 11205883.0, 0, 0}, T);
    return sgte({-21530.653, 222.396264, -31.4, 0, 0, 0, 0, 0}, T);
}

// Formation energy of Fe3B from 3 bcc-Fe + beta-B (GFA.cu:52).
__host__ __device__ inline double dgFormFe3B(double T) { return -77749.0 + 2.59 * T; }

// Inden's function f(tau) with structure factor p = 0.6 (GFA.cu:53-62); the
// decimals are 79/(140 p A), (474/497)(1/p - 1)/(6A, 135A, 600A) below
// tau = 1 and 1/(10A, 315A, 1500A) above, A = 518/1125 + (11692/15975)(1/p - 1).
__host__ __device__ inline double indenF(double tau) {
    if (tau <= 1.0) {
        const double t3 = tau * tau * tau, t9 = t3 * t3 * t3, t15 = t9 * t3 * t3;
        return 1.0 - 9.9167285e-1 / tau - 1.11737779e-1 * t3 - 4.96612349e-3 * t9 - 1.11737779e-3 * t15;
    }
    const double r = 1.0 / tau, r5 = r * r * r * r * r, r10 = r5 * r5, r15 = r10 * r5, r25 = r15 * r10;
    return -1.05443689e-1 * r5 - 3.34741816e-3 * r15 - 7.02957924e-4 * r25;
}

// Interpolation h(x) = x^3 (10 - 15x + 6x^2), its derivative, and the
// double-well derivative g'(x) = 2x(1 - x)(1 - 2x) (GFA.cu:64-66).
__host__ __device__ inline double hInterp(double x) { return x * x * x * (10.0 - 15.0 * x + 6.0 * x * x); }
__host__ __device__ inline double hInterpPrime(double x) { return 30.0 * x * x * (1.0 - x) * (1.0 - x); }
__host__ __device__ inline double wellPrime(double x) { return 2.0 * x * (1.0 - x) * (1.0 - 2.0 * x); }

// Mole fraction of B clamped to [1e-12, 1 - 1e-12] before the logarithms
// (GFA.cu:69,74).
__host__ __device__ inline double clampFraction(double c) { return fmax(1e-12, fmin(1.0 - 1e-12, c)); }

// dG_L/dc of the regular-solution liquid (GFA.cu:68-72), J/mol.
__host__ __device__ inline double dGLiquidDc(double c, double T) {
    const double x = clampFraction(c);
    return (gLiquidB(T) - gLiquidFe(T)) + kR * T * log(x / (1.0 - x)) + (1.0 - 2.0 * x) * lLiquidBFe(T);
}

// G_L(c) of the liquid (GFA.cu:73-78), J/mol.
__host__ __device__ inline double gLiquid(double c, double T) {
    const double x = clampFraction(c);
    return x * gLiquidB(T) + (1.0 - x) * gLiquidFe(T) + kR * T * (x * log(x) + (1.0 - x) * log(1.0 - x)) +
           x * (1.0 - x) * lLiquidBFe(T);
}

// Free-energy densities, J/m^3 (GFA.cu:79-85): liquid, the Fe3B solid, and
// the amorphous-minus-liquid difference R Tg ln(1 + alpha) f(T/Tg) / Vm_L.
__host__ __device__ inline double fLiquid(double c, double T, double vmL) { return gLiquid(c, T) / vmL; }
__host__ __device__ inline double fSolid(double T, double vmS) {
    return (3.0 * gBccFe(T) + gBetaB(T) + dgFormFe3B(T)) / vmS;
}
__host__ __device__ inline double dfAmorphous(double T, double Tg, double alpha, double vmL) {
    return kR * Tg * log(1.0 + alpha) * indenF(T / Tg) / vmL;
}

}  // namespace gfa
}  // namespace PhiX
