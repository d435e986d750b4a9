// gfa.hip -- one fused launch per Euler step of the Fe-B glass-formation
// model (the reference's GFA_2D program, applications/solvers/glass_formation/
// 2D/GFA.cu:143-201; its six multi-steps run in Solver.cu:317-338 order).
//
// Why one pass is exact: every multi-step reads its inputs either pointwise at
// the same cell or through a radius-1 stencil of fields whose ghost layer is
// set at the start of the step. mu is pointwise in (c, phi), and its BC fill
// (periodic wrap, zero-flux copy) equals evaluating it at the BC image of c
// and phi. D's ghost layer is zero for every BC: multi-step 3 fills it and
// then recomputes D as a STEADY unknown, and computeRHS clears the whole
// stored array before the pointwise term writes the physical cells
// (Solver.cu:322-326, Equation.cu:265, TermPW.inl:28-36). So a workgroup
// loads c, phi, eta on its tile plus a one-cell ring through the BC index
// maps, evaluates mu (and, inside the domain, D) there into LDS, and updates
// c (D lap mu + grad D . grad mu), then phi reading the new c, then eta
// reading the new phi, all at the same cell. HBM per cell-step: read c, phi,
// eta, write them (48 B in FP64); mu and D never leave the chip.
//
// FP64 throughout. Results equal the generic per-term path up to summation
// order, FMA contraction and the divisions by Vm_L and R T, done here as
// products with their reciprocals (the numpy oracle agrees to ~1e-15). The
// kernel is bound by FP64 arithmetic (about 3.3 logarithms per cell with the
// ring's share), not by HBM.
#include <cmath>
#include <utility>

#include "common.hpp"

#include "../apps/gfa_model.h"

namespace phx {
namespace {

constexpr int GTX = 32, GTY = 16, GNT = GTX * GTY;  // tile and threads
constexpr int GHX = GTX + 2, GHY = GTY + 2;         // tile plus one-cell ring

// Constants of one run: the settings' values and what depends only on T
// (GFA.cu:117-121; the SGTE energies at T, hoisted out of the per-cell work).
struct GfaConst {
    double dt, idx2, idy2, i2dx, i2dy;
    double gB, gFe, lBFe, RT, ivmL, iRT;  // J/mol at T; R T; 1 / Vm_L; 1 / (R T)
    double D_L, D_S, D_Am;               // diffusivities, m^2/s
    double fS, dfAm;                     // J/m^3
    double M_phi, k_phi, M_eta, k_eta;   // mobilities; mobility x gradient coefficient
    double w_phi, w_eta, w_ex;
};

struct GfaArgs {
    Layout L;
    AxisBC<double> bx, by;
    const double* c;
    const double* phi;
    const double* eta;
    double* co;
    double* po;
    double* eo;
    GfaConst k;
};

// mu = (1 - h(phi)) dG_L/dc / Vm_L  (GFA.cu:145-147, 68-72)
__device__ __forceinline__ double gfa_mu(const GfaConst& k, double c, double p) {
    const double x = PhiX::gfa::clampFraction(c);
    const double dg = (k.gB - k.gFe) + k.RT * log(x / (1.0 - x)) + (1.0 - 2.0 * x) * k.lBFe;
    return (1.0 - PhiX::gfa::hInterp(p)) * dg * k.ivmL;
}

// D = M_c(phi, eta) c (1 - c)  (GFA.cu:151-158)
__device__ __forceinline__ double gfa_D(const GfaConst& k, double c, double p, double e) {
    const double hp = PhiX::gfa::hInterp(p), he = PhiX::gfa::hInterp(e);
    const double Mc = ((1.0 - hp) * ((1.0 - he) * k.D_L + he * k.D_Am) + hp * k.D_S) * k.iRT;
    return Mc * c * (1.0 - c);
}

// f_L(c) = G_L(c) / Vm_L  (GFA.cu:73-79)
__device__ __forceinline__ double gfa_fL(const GfaConst& k, double c) {
    const double x = PhiX::gfa::clampFraction(c);
    const double g = x * k.gB + (1.0 - x) * k.gFe + k.RT * (x * log(x) + (1.0 - x) * log(1.0 - x)) +
                     x * (1.0 - x) * k.lBFe;
    return g * k.ivmL;
}

// Source cell of ghost layer l (1-based) beyond the low (hi = false) or high
// face of an axis of n cells: periodic wraps, zero-flux copies the face cell
// into every layer (src/boundary/Boundary.cu:276-329).
__device__ __forceinline__ int ghost_src(int l, bool hi, int n, int kind) {
    if (kind == PHX_BC_PERIODIC) return hi ? l - 1 : n - l;
    return hi ? n - 1 : 0;
}

__global__ __launch_bounds__(GNT) void k_gfa_step(GfaArgs A) {
    __shared__ double s_mu[GHY][GHX], s_D[GHY][GHX], s_p[GHY][GHX], s_e[GHY][GHX];
    const GfaConst& k = A.k;
    const int i0 = blockIdx.x * GTX, j0 = blockIdx.y * GTY;
    for (int q = threadIdx.x; q < GHX * GHY; q += GNT) {
        const int hx = q % GHX, hy = q / GHX;
        const long long o = A.L.at(map_in(i0 + hx - 1, A.bx), map_in(j0 + hy - 1, A.by), 0);
        const double c = A.c[o], p = A.phi[o], e = A.eta[o];
        const int x = i0 + hx - 1, y = j0 + hy - 1;
        const bool inside = x >= 0 && x < A.L.n0 && y >= 0 && y < A.L.n1;
        s_mu[hy][hx] = gfa_mu(k, c, p);
        s_D[hy][hx] = inside ? gfa_D(k, c, p, e) : 0.0;   // D's ghost layer is zero
        s_p[hy][hx] = p;
        s_e[hy][hx] = e;
    }
    __syncthreads();
    const int tx = threadIdx.x % GTX, ty = threadIdx.x / GTX;
    const int i = i0 + tx, j = j0 + ty;
    if (i >= A.L.n0 || j >= A.L.n1) return;
    const int X = tx + 1, Y = ty + 1;
    const long long o = A.L.at(i, j, 0);

    // step 4: c += dt (D lap mu + dD/dx dmu/dx + dD/dy dmu/dy)  (GFA.cu:160-165)
    const double muC = s_mu[Y][X], muE = s_mu[Y][X + 1], muW = s_mu[Y][X - 1];
    const double muN = s_mu[Y + 1][X], muS = s_mu[Y - 1][X];
    const double lap_mu = (muE - 2.0 * muC + muW) * k.idx2 + (muN - 2.0 * muC + muS) * k.idy2;
    double rc = s_D[Y][X] * lap_mu;
    rc += ((s_D[Y][X + 1] - s_D[Y][X - 1]) * k.i2dx) * ((muE - muW) * k.i2dx);
    rc += ((s_D[Y + 1][X] - s_D[Y - 1][X]) * k.i2dy) * ((muN - muS) * k.i2dy);
    const double c0 = A.c[o];
    const double cn = c0 + k.dt * rc;

    // step 5: phi, reading the new c  (GFA.cu:167-178)
    const double p0 = s_p[Y][X], e0 = s_e[Y][X];
    const double lap_p = (s_p[Y][X + 1] - 2.0 * p0 + s_p[Y][X - 1]) * k.idx2 +
                         (s_p[Y + 1][X] - 2.0 * p0 + s_p[Y - 1][X]) * k.idy2;
    const double dFp = PhiX::gfa::hInterpPrime(p0) * (k.fS - gfa_fL(k, cn) - PhiX::gfa::hInterp(e0) * k.dfAm) +
                       k.w_phi * PhiX::gfa::wellPrime(p0) + 2.0 * k.w_ex * p0 * e0 * e0;
    const double pn = p0 + k.dt * (-k.M_phi * dFp + k.k_phi * lap_p);

    // step 6: eta, reading the new phi  (GFA.cu:180-190)
    const double lap_e = (s_e[Y][X + 1] - 2.0 * e0 + s_e[Y][X - 1]) * k.idx2 +
                         (s_e[Y + 1][X] - 2.0 * e0 + s_e[Y - 1][X]) * k.idy2;
    const double dFe = (1.0 - PhiX::gfa::hInterp(pn)) * PhiX::gfa::hInterpPrime(e0) * k.dfAm +
                       k.w_eta * PhiX::gfa::wellPrime(e0) + 2.0 * k.w_ex * pn * pn * e0;
    const double en = e0 + k.dt * (-k.M_eta * dFe + k.k_eta * lap_e);

    A.co[o] = cn;
    A.po[o] = pn;
    A.eo[o] = en;

    // Ghost cells of the output: the BC fill of this step's input state, as
    // the reference's multi-step leaves them (BCs on each source field before
    // its update; the update adds nothing on ghosts). Face cells write the
    // ghost layers beyond them; corner cells also the corner blocks.
    const int n0 = A.L.n0, n1 = A.L.n1, g = A.L.g;
    const bool xlo = i == 0, xhi = i == n0 - 1, ylo = j == 0, yhi = j == n1 - 1;
    if (!(xlo || xhi || ylo || yhi)) return;
    auto put = [&](int gi, int gj, int si, int sj) {
        const long long d = A.L.at(gi, gj, 0), s = A.L.at(si, sj, 0);
        A.co[d] = A.c[s];
        A.po[d] = A.phi[s];
        A.eo[d] = A.eta[s];
    };
    for (int l = 1; l <= g; ++l) {
        if (xlo) put(-l, j, ghost_src(l, false, n0, A.bx.klo), j);
        if (xhi) put(n0 - 1 + l, j, ghost_src(l, true, n0, A.bx.khi), j);
        if (ylo) put(i, -l, i, ghost_src(l, false, n1, A.by.klo));
        if (yhi) put(i, n1 - 1 + l, i, ghost_src(l, true, n1, A.by.khi));
        for (int m = 1; m <= g; ++m) {
            if (xlo && ylo) put(-l, -m, ghost_src(l, false, n0, A.bx.klo), ghost_src(m, false, n1, A.by.klo));
            if (xhi && ylo) put(n0 - 1 + l, -m, ghost_src(l, true, n0, A.bx.khi), ghost_src(m, false, n1, A.by.klo));
            if (xlo && yhi) put(-l, n1 - 1 + m, ghost_src(l, false, n0, A.bx.klo), ghost_src(m, true, n1, A.by.khi));
            if (xhi && yhi)
                put(n0 - 1 + l, n1 - 1 + m, ghost_src(l, true, n0, A.bx.khi), ghost_src(m, true, n1, A.by.khi));
        }
    }
}

bool params_ok(const phx_gfa_params* p) {
    const double v[] = {p->dt, p->M_eta, p->eps_sq, p->beta_sq, p->w_phi, p->w_eta, p->w_ex,
                        p->T, p->T_g, p->alpha, p->V_m_L, p->V_m_S, p->D_Am};
    for (double x : v)
        if (!std::isfinite(x)) return false;
    return p->T > 0.0 && p->T_g > 0.0 && p->alpha > -1.0 && p->V_m_L != 0.0 && p->V_m_S != 0.0;
}

}  // namespace
}  // namespace phx

using namespace phx;

extern "C" {

size_t phx_gfa_workspace_bytes(const phx_grid* G) {
    if (!G || check_grid(G, "phx_gfa_workspace_bytes") != PHX_OK) return 0;
    return 3 * (size_t)make_layout(G).total() * sizeof(double);
}

int phx_gfa_euler_steps(const phx_grid* G, double* c, double* phi, double* eta, void* work,
                        const phx_gfa_params* p, int nsteps, phx_stream stream) {
    if (!G || !c || !phi || !eta || !p) return fail(PHX_ERR_INVALID, "phx_gfa_euler_steps: null argument");
    if (check_grid_fused(G, "phx_gfa_euler_steps") != PHX_OK) return PHX_ERR_INVALID;
    if (G->dim != 2) return fail(PHX_ERR_INVALID, "phx_gfa_euler_steps: the GFA model is two-dimensional (dim 2)");
    for (int a = 0; a < 2; ++a)
        for (int s = 0; s < 2; ++s)
            if (G->bc[a][s] != PHX_BC_PERIODIC && G->bc[a][s] != PHX_BC_NOFLUX)
                return fail(PHX_ERR_INVALID, "phx_gfa_euler_steps: only periodic and zero-flux faces (as buildBCs)");
    if (G->ghost > G->n[0] || G->ghost > G->n[1])
        return fail(PHX_ERR_INVALID, "phx_gfa_euler_steps: ghost layers must not exceed the axis size");
    if (nsteps < 0) return fail(PHX_ERR_INVALID, "phx_gfa_euler_steps: nsteps < 0");
    if (!params_ok(p)) return fail(PHX_ERR_INVALID, "phx_gfa_euler_steps: non-finite or out-of-range constant");
    if (nsteps == 0) return PHX_OK;
    if (!work) return fail(PHX_ERR_INVALID, "phx_gfa_euler_steps: null work (phx_gfa_workspace_bytes)");

    using namespace PhiX::gfa;
    const double T = p->T;
    GfaArgs A{};
    A.L = make_layout(G);
    A.bx = AxisBC<double>{G->n[0], G->bc[0][0], G->bc[0][1], 0.0, 0.0};
    A.by = AxisBC<double>{G->n[1], G->bc[1][0], G->bc[1][1], 0.0, 0.0};
    GfaConst& k = A.k;
    k.dt = p->dt;
    k.idx2 = 1.0 / (G->d[0] * G->d[0]);
    k.idy2 = 1.0 / (G->d[1] * G->d[1]);
    k.i2dx = 0.5 / G->d[0];
    k.i2dy = 0.5 / G->d[1];
    k.gB = gLiquidB(T);
    k.gFe = gLiquidFe(T);
    k.lBFe = lLiquidBFe(T);
    k.RT = kR * T;
    k.ivmL = 1.0 / p->V_m_L;
    k.iRT = 1.0 / (kR * T);
    k.D_L = 2.0e-6 * exp(-1.11e5 / (kR * T));
    k.D_S = 1.311e-6 * exp(-1.51e5 / (kR * T));
    k.D_Am = p->D_Am;
    k.fS = fSolid(T, p->V_m_S);
    k.dfAm = dfAmorphous(T, p->T_g, p->alpha, p->V_m_L);
    k.M_phi = 22.1 * exp(-140.0e3 / (kR * T));
    k.k_phi = k.M_phi * p->eps_sq;
    k.M_eta = p->M_eta;
    k.k_eta = p->M_eta * p->beta_sq;
    k.w_phi = p->w_phi;
    k.w_eta = p->w_eta;
    k.w_ex = p->w_ex;

    hipStream_t st = (hipStream_t)stream;
    const long long N = A.L.total();
    double* w = (double*)work;
    double* in[3] = {c, phi, eta};
    double* out[3] = {w, w + N, w + 2 * N};
    const dim3 grid(cdiv(G->n[0], GTX), cdiv(G->n[1], GTY));
    for (int s = 0; s < nsteps; ++s) {
        A.c = in[0];
        A.phi = in[1];
        A.eta = in[2];
        A.co = out[0];
        A.po = out[1];
        A.eo = out[2];
        launch_timer_begin(st);
        hipLaunchKernelGGL(k_gfa_step, grid, dim3(GNT), 0, st, A);
        launch_timer_end(10, st);
        PHX_HIP(hipGetLastError());
        for (int f = 0; f < 3; ++f) std::swap(in[f], out[f]);
    }
    if (nsteps % 2) {
        // the state is in the workspace: copy its z = 0 slab (physical cells and
        // the x/y ghost ring) back; the z ghost planes of the caller's arrays stay
        const long long off = A.L.sxy * A.L.g;
        double* dst[3] = {c, phi, eta};
        for (int f = 0; f < 3; ++f)
            PHX_HIP(hipMemcpyAsync(dst[f] + off, in[f] + off, sizeof(double) * A.L.sxy, hipMemcpyDeviceToDevice, st));
    }
    return PHX_OK;
}

}  // extern "C"
