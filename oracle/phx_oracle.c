/*
 * phx_oracle.c -- FP64 CPU restatement of PhiX's explicit finite-difference
 * phase-field step. TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke
 * check and by the CPU-baseline leg of bench.py, always as the checker, never as
 * the product path. Built HIP-free with gcc -ffp-contract=off so every FP64
 * operation happens in the order written below, which is the order of the
 * reference source cited at each function.
 *
 * Layout (include/field/ScalarField.h:14-26, 90-96): storedDims[a] = n[a]+2g for
 * all three axes; index = (i+g) + sx*((j+g) + sy*(k+g)).
 *
 * Pins (what checks each function independently of this restatement; the
 * tests named live in tests/):
 *   orc_bc_apply / orc_bc_fill      exhaustive ghost-source index checks on tiny
 *                                   grids against a numpy model written from the
 *                                   BC definitions (test_oracle_bc.py)
 *   orc_lap_accumulate              closed-form Fourier eigenvalue of the
 *                                   periodic 5/7-point Laplacian (test_oracle_stencil.py)
 *   orc_grad_accumulate             exact on linear/periodic-sine fields
 *   orc_pw_dw_accumulate,
 *   orc_mul_accumulate              test_fieldops known answers (develop/fieldOperation/
 *                                   test_fieldops.cu:118-291) and f'(c) roots
 *   orc_ch_euler_steps,
 *   orc_ch_rk4_steps,
 *   orc_ac_rk4_steps                per-mode amplification factors 1+z and
 *                                   1+z+z^2/2+z^3/6+z^4/24 for linearised f'
 *                                   (rho=0 / single-mode), mass conservation,
 *                                   energy decay (test_oracle_steppers.py)
 *   orc_mass_energy                 math.fsum of the same per-cell terms and a
 *                                   numpy roll formula (test_oracle_reduce.py)
 */
#include "../include/phix_amd.h"

#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* The per-cell passes run on OpenMP threads (each cell writes only its own
 * output, so the results do not depend on the thread count); orc_set_threads
 * fixes the count (1 = the reference's serial loops). Reductions stay serial. */
#ifdef _OPENMP
#define PAR_KJ _Pragma("omp parallel for collapse(2) schedule(static)")
#define PAR_Q _Pragma("omp parallel for schedule(static)")
#else
#define PAR_KJ
#define PAR_Q
#endif

int orc_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

/* ----------------------------------------------------------------------- */
/* layout                                                                  */
/* ----------------------------------------------------------------------- */
typedef struct {
    int n[3], g, s[3];
    long sx, sxy, total;
} lay_t;

static lay_t mklay(const phx_grid* G) {
    lay_t L;
    for (int a = 0; a < 3; ++a) { L.n[a] = G->n[a]; L.s[a] = G->n[a] + 2 * G->ghost; }
    L.g = G->ghost;
    L.sx = L.s[0];
    L.sxy = (long)L.s[0] * L.s[1];
    L.total = L.sxy * L.s[2];
    return L;
}

static inline long ix(const lay_t* L, int i, int j, int k) {
    return (long)(i + L->g) + L->sx * (long)(j + L->g) + L->sxy * (long)(k + L->g);
}

size_t orc_stored_size(const phx_grid* G) { return (size_t)mklay(G).total; }

/* ----------------------------------------------------------------------- */
/* boundary conditions: src/boundary/Boundary.cu:43-84 (face geometry),    */
/* :204-222 periodic, :248-274 no-flux, :305-329 fixed. The face loop runs */
/* over the full stored extent of the other two axes.                      */
/* ----------------------------------------------------------------------- */
void orc_bc_apply(const phx_grid* G, double* f, int axis, int kind, int side, double value) {
    lay_t L = mklay(G);
    long astride, fs0, fs1;
    int nf0, nf1, na = L.n[axis], g = L.g;
    if (axis == 0) { astride = 1;     nf0 = L.s[1]; nf1 = L.s[2]; fs0 = L.sx; fs1 = L.sxy; }
    else if (axis == 1) { astride = L.sx; nf0 = L.s[0]; nf1 = L.s[2]; fs0 = 1;    fs1 = L.sxy; }
    else { astride = L.sxy; nf0 = L.s[0]; nf1 = L.s[1]; fs0 = 1; fs1 = L.sx; }
    int lo = (side == PHX_SIDE_LOW || side == PHX_SIDE_BOTH);
    int hi = (side == PHX_SIDE_HIGH || side == PHX_SIDE_BOTH);
    for (int t0 = 0; t0 < nf0; ++t0)
        for (int t1 = 0; t1 < nf1; ++t1) {
            long off = t0 * fs0 + t1 * fs1;
            if (kind == PHX_BC_PERIODIC) {
                for (int l = 1; l <= g; ++l) {
                    f[(g - l) * astride + off] = f[(g + na - l) * astride + off];
                    f[(g + na + l - 1) * astride + off] = f[(g + l - 1) * astride + off];
                }
            } else if (kind == PHX_BC_NOFLUX) {
                if (lo) for (int l = 1; l <= g; ++l) f[(g - l) * astride + off] = f[g * astride + off];
                if (hi) for (int l = 1; l <= g; ++l)
                        f[(g + na + l - 1) * astride + off] = f[(g + na - 1) * astride + off];
            } else if (kind == PHX_BC_FIXED) {
                if (lo) for (int l = 1; l <= g; ++l) f[(g - l) * astride + off] = value;
                if (hi) for (int l = 1; l <= g; ++l) f[(g + na + l - 1) * astride + off] = value;
            }
        }
}

/* The BC list {X, Y, Z} of the active axes, as apps hand it to the Solver.
 * A periodic axis is one PeriodicBC object; otherwise one object per side. */
void orc_bc_fill(const phx_grid* G, double* f) {
    for (int a = 0; a < G->dim; ++a) {
        int klo = G->bc[a][0], khi = G->bc[a][1];
        if (klo == PHX_BC_PERIODIC && khi == PHX_BC_PERIODIC) {
            orc_bc_apply(G, f, a, PHX_BC_PERIODIC, PHX_SIDE_BOTH, 0.0);
            continue;
        }
        if (klo != PHX_BC_NONE) orc_bc_apply(G, f, a, klo, PHX_SIDE_LOW, G->bc_value[a][0]);
        if (khi != PHX_BC_NONE) orc_bc_apply(G, f, a, khi, PHX_SIDE_HIGH, G->bc_value[a][1]);
    }
}

/* ----------------------------------------------------------------------- */
/* term kernels (CPU twins of the reference term launchers)                */
/* ----------------------------------------------------------------------- */

/* rhs += coeff*lap(src): src/equation/Equation.cu:164-180 */
void orc_lap_accumulate(const phx_grid* G, double* rhs, const double* src, double coeff) {
    lay_t L = mklay(G);
    int dim = G->dim;
    double idx2 = 1.0 / (G->d[0] * G->d[0]);
    double idy2 = (dim >= 2) ? 1.0 / (G->d[1] * G->d[1]) : 0.0;
    double idz2 = (dim >= 3) ? 1.0 / (G->d[2] * G->d[2]) : 0.0;
    PAR_KJ
    for (int k = 0; k < L.n[2]; ++k)
        for (int j = 0; j < L.n[1]; ++j)
            for (int i = 0; i < L.n[0]; ++i) {
                long c = ix(&L, i, j, k);
                double val = (src[c + 1] - 2.0 * src[c] + src[c - 1]) * idx2;
                if (dim >= 2) val += (src[c + L.sx] - 2.0 * src[c] + src[c - L.sx]) * idy2;
                if (dim >= 3) val += (src[c + L.sxy] - 2.0 * src[c] + src[c - L.sxy]) * idz2;
                rhs[c] += coeff * val;
            }
}

/* rhs += coeff*(src[+1]-src[-1])*inv_2d along axis: Equation.cu:212-237 */
void orc_grad_accumulate(const phx_grid* G, double* rhs, const double* src, int axis, double coeff) {
    lay_t L = mklay(G);
    double inv_2d = 0.5 / G->d[axis];
    long st = axis == 0 ? 1 : (axis == 1 ? L.sx : L.sxy);
    PAR_KJ
    for (int k = 0; k < L.n[2]; ++k)
        for (int j = 0; j < L.n[1]; ++j)
            for (int i = 0; i < L.n[0]; ++i) {
                long c = ix(&L, i, j, k);
                rhs[c] += coeff * (src[c + st] - src[c - st]) * inv_2d;
            }
}

/* rhs += coeff*s1*s2: Equation.cu:410-420 */
void orc_mul_accumulate(const phx_grid* G, double* rhs, const double* s1, const double* s2, double coeff) {
    lay_t L = mklay(G);
    PAR_KJ
    for (int k = 0; k < L.n[2]; ++k)
        for (int j = 0; j < L.n[1]; ++j)
            for (int i = 0; i < L.n[0]; ++i) {
                long c = ix(&L, i, j, k);
                rhs[c] += coeff * s1[c] * s2[c];
            }
}

/* The double-well derivative exactly as the CH app writes it
 * (applications/solvers/Cahn-Hillard_double-well/2D/Cahn-Hillard_double-well.cu:87-90). */
double orc_dw(double c, double rho, double ca, double cb) {
    return 2.0 * rho * (c - ca) * (c - cb) * (2.0 * c - ca - cb);
}

/* rhs += coeff*F(src): include/equation/TermPW.inl:129-138 */
void orc_pw_dw_accumulate(const phx_grid* G, double* rhs, const double* src,
                          double rho, double ca, double cb, double coeff) {
    lay_t L = mklay(G);
    PAR_KJ
    for (int k = 0; k < L.n[2]; ++k)
        for (int j = 0; j < L.n[1]; ++j)
            for (int i = 0; i < L.n[0]; ++i) {
                long c = ix(&L, i, j, k);
                rhs[c] += coeff * orc_dw(src[c], rho, ca, cb);
            }
}

/* ----------------------------------------------------------------------- */
/* Cahn-Hilliard                                                           */
/* ----------------------------------------------------------------------- */

/* mu = pw(c, f') - kappa*lap(c): Equation::computeRHSCPU (Equation.cu:278-291)
 * zeroes the whole stored array, then accumulates the terms in order. */
void orc_ch_mu(const phx_grid* G, const phx_ch_params* p, const double* c, double* mu) {
    lay_t L = mklay(G);
    memset(mu, 0, sizeof(double) * L.total);
    orc_pw_dw_accumulate(G, mu, c, p->rho, p->ca, p->cb, 1.0);
    orc_lap_accumulate(G, mu, c, -p->kappa);
}

/* rhs = M*lap(mu(u), bcs): computeRHSCPU with the single composite term of
 * Equation.cu:578-591 (materialise, BCs on a temporary, stencil). */
void orc_ch_rhs(const phx_grid* G, const phx_ch_params* p, const double* u, double* mu, double* rhs) {
    lay_t L = mklay(G);
    memset(rhs, 0, sizeof(double) * L.total);
    orc_ch_mu(G, p, u, mu);
    orc_bc_fill(G, mu);
    orc_lap_accumulate(G, rhs, mu, p->M);
}

/* Multi-step Euler (Solver.cu:342-360) for
 *   {c, bcs, mu = f'(c)-kappa lap c, STEADY}, {mu, bcs, dc/dt = M lap mu, TRANSIENT}.
 * scratch: one stored field. */
void orc_ch_euler_steps(const phx_grid* G, const phx_ch_params* p, double* c, double* mu,
                        double* scratch, int nsteps) {
    lay_t L = mklay(G);
    for (int s = 0; s < nsteps; ++s) {
        orc_bc_fill(G, c);
        orc_ch_mu(G, p, c, mu);
        orc_bc_fill(G, mu);
        memset(scratch, 0, sizeof(double) * L.total);
        orc_lap_accumulate(G, scratch, mu, p->M);
        PAR_Q
        for (long q = 0; q < L.total; ++q) c[q] += p->dt * scratch[q];
    }
}

/* Work: 6 stored fields (k1..k4, tmp, mu). */
size_t orc_ch_rk4_work_doubles(const phx_grid* G) { return 6 * (size_t)mklay(G).total; }

/* Classic RK4 with BCs on each stage state -- the GPU semantics of
 * Solver::rk4AdvanceGPU (Solver.cu:192-259), not the CPU twin whose stages 2-4
 * read stale ghosts (Solver.cu:281-283). */
void orc_ch_rk4_steps(const phx_grid* G, const phx_ch_params* p, double* phi, double* work, int nsteps) {
    lay_t L = mklay(G);
    long N = L.total;
    double *k1 = work, *k2 = work + N, *k3 = work + 2 * N, *k4 = work + 3 * N;
    double *tmp = work + 4 * N, *mu = work + 5 * N;
    double dt = p->dt, dt2 = dt * 0.5, dt6 = dt / 6.0;
    for (int s = 0; s < nsteps; ++s) {
        orc_bc_fill(G, phi);
        orc_ch_rhs(G, p, phi, mu, k1);
        PAR_Q
        for (long q = 0; q < N; ++q) tmp[q] = phi[q] + dt2 * k1[q];
        orc_bc_fill(G, tmp);
        orc_ch_rhs(G, p, tmp, mu, k2);
        PAR_Q
        for (long q = 0; q < N; ++q) tmp[q] = phi[q] + dt2 * k2[q];
        orc_bc_fill(G, tmp);
        orc_ch_rhs(G, p, tmp, mu, k3);
        PAR_Q
        for (long q = 0; q < N; ++q) tmp[q] = phi[q] + dt * k3[q];
        orc_bc_fill(G, tmp);
        orc_ch_rhs(G, p, tmp, mu, k4);
        PAR_Q
        for (long q = 0; q < N; ++q) phi[q] += dt6 * (k1[q] + 2.0 * k2[q] + 2.0 * k3[q] + k4[q]);
    }
}

/* ----------------------------------------------------------------------- */
/* Allen-Cahn, N coupled order parameters (VectorSolver RK4, GPU semantics: */
/* src/solver/VectorSolver.cu:149-228)                                     */
/* ----------------------------------------------------------------------- */

/* rhs_i = (L*kappa)*lap(u_i) + pw(u_i, e-e^3, L) + sum_{j!=i} pw(u_i,u_j, a*b*b, -2*L*gamma) */
static void orc_ac_rhs(const phx_grid* G, const phx_ac_params* p, int nf, double* const* u, double** rhs) {
    lay_t L = mklay(G);
    double cl = p->L * p->kappa, cp = p->L, cg = -2.0 * p->L * p->gamma;
    for (int f = 0; f < nf; ++f) {
        memset(rhs[f], 0, sizeof(double) * L.total);
        orc_lap_accumulate(G, rhs[f], u[f], cl);
        for (int k = 0; k < L.n[2]; ++k)
            for (int j = 0; j < L.n[1]; ++j)
                for (int i = 0; i < L.n[0]; ++i) {
                    long c = ix(&L, i, j, k);
                    double e = u[f][c];
                    rhs[f][c] += cp * (e - e * e * e);
                }
        for (int h = 0; h < nf; ++h) {
            if (h == f) continue;
            for (int k = 0; k < L.n[2]; ++k)
                for (int j = 0; j < L.n[1]; ++j)
                    for (int i = 0; i < L.n[0]; ++i) {
                        long c = ix(&L, i, j, k);
                        double a = u[f][c], b = u[h][c];
                        rhs[f][c] += cg * (a * b * b);
                    }
        }
    }
}

/* work: 5*nf stored fields */
void orc_ac_rk4_steps(const phx_grid* G, const phx_ac_params* p, int nf, double** eta, double* work, int nsteps) {
    lay_t L = mklay(G);
    long N = L.total;
    double* k[4][16];
    double* tmp[16];
    for (int f = 0; f < nf; ++f) {
        for (int s = 0; s < 4; ++s) k[s][f] = work + (long)(s * nf + f) * N;
        tmp[f] = work + (long)(4 * nf + f) * N;
    }
    double dt = p->dt, dt2 = dt * 0.5, dt6 = dt / 6.0;
    double a[3] = {dt2, dt2, dt};
    for (int s = 0; s < nsteps; ++s) {
        for (int f = 0; f < nf; ++f) orc_bc_fill(G, eta[f]);
        orc_ac_rhs(G, p, nf, eta, k[0]);
        for (int st = 1; st < 4; ++st) {
            for (int f = 0; f < nf; ++f)
                for (long q = 0; q < N; ++q) tmp[f][q] = eta[f][q] + a[st - 1] * k[st - 1][f][q];
            for (int f = 0; f < nf; ++f) orc_bc_fill(G, tmp[f]);
            orc_ac_rhs(G, p, nf, tmp, k[st]);
        }
        for (int f = 0; f < nf; ++f)
            for (long q = 0; q < N; ++q)
                eta[f][q] += dt6 * (k[0][f][q] + 2.0 * k[1][f][q] + 2.0 * k[2][f][q] + k[3][f][q]);
    }
}

/* ----------------------------------------------------------------------- */
/* FP64 diagnostics                                                        */
/* ----------------------------------------------------------------------- */

/* Neumaier's improvement of Kahan summation (Neumaier 1974). */
typedef struct { double s, c; } nsum_t;
static inline void nadd(nsum_t* a, double x) {
    double t = a->s + x;
    if (fabs(a->s) >= fabs(x)) a->c += (a->s - t) + x;
    else a->c += (x - t) + a->s;
    a->s = t;
}

/* Neighbour value along one axis with the grid BC table (physical index q, step +-1). */
static double nb(const phx_grid* G, const lay_t* L, const double* c, int q[3], int axis, int step) {
    int r[3] = {q[0], q[1], q[2]};
    int p = q[axis] + step, n = L->n[axis];
    if (p < 0 || p >= n) {
        int side = p < 0 ? 0 : 1, kind = G->bc[axis][side];
        if (kind == PHX_BC_FIXED) return G->bc_value[axis][side];
        if (kind == PHX_BC_PERIODIC) p = (p + n) % n;
        else p = p < 0 ? 0 : n - 1;     /* NOFLUX (zero-gradient ghost) */
    }
    r[axis] = p;
    return c[ix(L, r[0], r[1], r[2])];
}

/* Per-cell free-energy density of computeFreeEnergy.py:76-98 (the numpy roll
 * is the periodic case of the neighbour rule above). */
double orc_energy_density(const phx_grid* G, const phx_ch_params* p, const double* c, int i, int j, int k) {
    lay_t L = mklay(G);
    int q[3] = {i, j, k};
    double v = c[ix(&L, i, j, k)];
    double fchem = p->rho * (v - p->ca) * (v - p->ca) * (p->cb - v) * (p->cb - v);
    double gs = 0.0;
    for (int a = 0; a < G->dim; ++a) {
        double dd = (nb(G, &L, c, q, a, +1) - nb(G, &L, c, q, a, -1)) / (2.0 * G->d[a]);
        gs += dd * dd;
    }
    return fchem + 0.5 * p->kappa * gs;
}

/* out[0] = sum c, out[1] = total energy = (sum density)*dV. Compensated sums. */
void orc_mass_energy(const phx_grid* G, const phx_ch_params* p, const double* c, double out[2]) {
    lay_t L = mklay(G);
    nsum_t m = {0, 0}, e = {0, 0};
    double dV = 1.0;
    for (int a = 0; a < G->dim; ++a) dV *= G->d[a];
    for (int k = 0; k < L.n[2]; ++k)
        for (int j = 0; j < L.n[1]; ++j)
            for (int i = 0; i < L.n[0]; ++i) {
                nadd(&m, c[ix(&L, i, j, k)]);
                nadd(&e, orc_energy_density(G, p, c, i, j, k));
            }
    out[0] = m.s + m.c;
    out[1] = (e.s + e.c) * dV;
}

/* ----------------------------------------------------------------------- */
/* SplitMix64 (Steele, Lea, Flood, "Fast splittable pseudorandom number    */
/* generators", OOPSLA 2014; constants of Vigna's public-domain splitmix64.c)*/
/* counter form: x_i = mix(seed + (i+1)*golden).                           */
/* ----------------------------------------------------------------------- */
static inline uint64_t sm_mix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
uint64_t orc_substream(uint64_t seed, uint64_t tag) {
    return sm_mix(seed ^ (0xD1B54A32D192ED03ULL * (tag + 1)));
}
/* out[q] = a + (b-a)*u_{offset+q},  u = (x >> 11) * 2^-53 */
void orc_uniform(uint64_t seed, uint64_t offset, size_t n, double a, double b, double* out) {
    for (size_t q = 0; q < n; ++q) {
        uint64_t x = sm_mix(seed + (offset + q + 1) * 0x9E3779B97F4A7C15ULL);
        double u = (double)(x >> 11) * (1.0 / 9007199254740992.0);
        out[q] = a + (b - a) * u;
    }
}
