// TermPW.inl -- pw() factories (included by Term.h; needs hipcc).
//
// Device side: one thread per physical cell, rhs[idx] += coeff*func(src[idx])
// (the term kernels of include/equation/TermPW.inl:18-90 in the reference),
// launched as a (x-blocks of 64, y, z) grid so no index division is needed.
//
// Host side: besides the launchers, pw() probes the functor on the host (a
// PHIX_FN lambda is __host__ __device__) and records in Term::sym whether it
// is, bit for bit at a set of probe points, one of the forms the fused
// kernels implement:
//   1 field : 2*rho*(c-ca)*(c-cb)*(2c-ca-cb)   (double-well derivative)
//             e - e*e*e                         (Allen-Cahn well)
//   2 fields: a*b*b                             (Allen-Cahn coupling)
// Anything else stays on the generic per-term path.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <stdexcept>
#include <string>

namespace PhiX {

namespace detail {

struct PwGeom {
    int nx, ny, nz, sx, sy, g;
};

inline PwGeom pwGeom(const ScalarField& f) {
    return PwGeom{f.mesh.n[0], f.mesh.n[1], f.mesh.n[2], f.storedDims[0], f.storedDims[1], f.ghost};
}

__host__ __device__ inline long long pwIndex(const PwGeom& G, int i, int j, int k) {
    return (long long)(i + G.g) + (long long)G.sx * ((long long)(j + G.g) + (long long)G.sy * (k + G.g));
}

inline dim3 pwGrid(const PwGeom& G) { return dim3((G.nx + 63) / 64, G.ny, G.nz); }

template <typename F>
__global__ void pw1Kernel(double* rhs, const double* a, F func, double coeff, PwGeom G) {
    int i = blockIdx.x * 64 + threadIdx.x;
    if (i >= G.nx) return;
    long long q = pwIndex(G, i, blockIdx.y, blockIdx.z);
    rhs[q] += coeff * func(a[q]);
}

template <typename F>
__global__ void pw2Kernel(double* rhs, const double* a, const double* b, F func, double coeff, PwGeom G) {
    int i = blockIdx.x * 64 + threadIdx.x;
    if (i >= G.nx) return;
    long long q = pwIndex(G, i, blockIdx.y, blockIdx.z);
    rhs[q] += coeff * func(a[q], b[q]);
}

template <typename F>
__global__ void pw3Kernel(double* rhs, const double* a, const double* b, const double* c, F func, double coeff,
                          PwGeom G) {
    int i = blockIdx.x * 64 + threadIdx.x;
    if (i >= G.nx) return;
    long long q = pwIndex(G, i, blockIdx.y, blockIdx.z);
    rhs[q] += coeff * func(a[q], b[q], c[q]);
}

void checkLaunch(const char* what);  // throws std::runtime_error on a launch error

// Probe points: generic values plus the wells' special points.
inline const double* probePoints(int& n) {
    static const double pts[] = {-2.1, -1.0, -0.37, 0.0, 0.11, 0.25, 0.3, 0.5, 0.7, 0.73, 0.999, 1.0, 1.9, 3.3};
    n = sizeof(pts) / sizeof(pts[0]);
    return pts;
}

inline double dwEval(double c, double rho, double ca, double cb) {
    return 2.0 * rho * (c - ca) * (c - cb) * (2.0 * c - ca - cb);
}

inline double snapDecimal(double x, double scale) { return std::round(x * scale) / scale; }

// Recovers (rho, ca, cb) of a double-well derivative from samples: the cubic
// through F(-1..2) has leading coefficient 4*rho, its inflection point is the
// midpoint m of the wells and F'(m) = -4*rho*h^2 with h the half distance.
// Accepts only if some candidate (raw, or snapped to 6/9/12 decimals, either
// well order) reproduces F bit for bit at every probe point.
template <typename F>
bool probeDoubleWell(F& func, double out[3]) {
    double v[4];
    for (int t = 0; t < 4; ++t) {
        v[t] = func(double(t - 1));
        if (!std::isfinite(v[t])) return false;
    }
    // Newton divided differences on x = -1, 0, 1, 2
    double d1 = v[1] - v[0], d2 = v[2] - v[1], d3 = v[3] - v[2];
    double e1 = (d2 - d1) / 2.0, e2 = (d3 - d2) / 2.0;
    double a3 = (e2 - e1) / 3.0;
    if (!(std::fabs(a3) > 0.0)) return false;
    // p(x) = v0 + d1 (x+1) + e1 (x+1) x + a3 (x+1) x (x-1)  ->  monomial coefficients
    double a2 = e1;
    double a1 = d1 + e1 - a3;
    double m = -a2 / (3.0 * a3);
    double dpm = a1 + 2.0 * a2 * m + 3.0 * a3 * m * m;
    double h2 = -dpm / a3;
    if (!(h2 > 0.0)) return false;
    double h = std::sqrt(h2);
    double raw[3] = {a3 / 4.0, m - h, m + h};
    const double scales[] = {0.0, 1e6, 1e9, 1e12};
    int np = 0;
    const double* pts = probePoints(np);
    for (double sc : scales) {
        double p[3];
        for (int q = 0; q < 3; ++q) p[q] = sc > 0 ? snapDecimal(raw[q], sc) : raw[q];
        for (int order = 0; order < 2; ++order) {
            double ca = order ? p[2] : p[1], cb = order ? p[1] : p[2];
            bool ok = true;
            for (int t = 0; t < np && ok; ++t) ok = (dwEval(pts[t], p[0], ca, cb) == func(pts[t]));
            if (ok) {
                out[0] = p[0];
                out[1] = ca;
                out[2] = cb;
                return true;
            }
        }
    }
    return false;
}

template <typename F>
bool probeAcWell(F& func) {
    int np = 0;
    const double* pts = probePoints(np);
    for (int t = 0; t < np; ++t) {
        double e = pts[t];
        if (!(func(e) == e - e * e * e)) return false;
    }
    return true;
}

template <typename F>
bool probeAbb(F& func) {
    int np = 0;
    const double* pts = probePoints(np);
    for (int t = 0; t < np; ++t)
        for (int s = 0; s < np; s += 3) {
            double a = pts[t], b = pts[s];
            if (!(func(a, b) == a * b * b)) return false;
        }
    return true;
}

inline void requireSameShape(const ScalarField& a, const ScalarField& b, const char* what) {
    if (a.mesh.n[0] != b.mesh.n[0] || a.mesh.n[1] != b.mesh.n[1] || a.mesh.n[2] != b.mesh.n[2] ||
        a.ghost != b.ghost)
        throw std::invalid_argument(std::string(what) + ": fields must share the same mesh dimensions and ghost width");
}

}  // namespace detail

template <typename Functor>
Term pw(const ScalarField& f, Functor func, double coeff) {
    Term t;
    t.type = TermType::POINTWISE;
    t.field = &f;
    t.coeff = coeff;
    t.sym.kind = TermSym::PW1;
    t.sym.f[0] = &f;
    t.sym.doublewell = detail::probeDoubleWell(func, t.sym.dw);
    t.sym.ac_well = detail::probeAcWell(func);
    const ScalarField* pf = &f;
    const detail::PwGeom G = detail::pwGeom(f);
    t.gpu_launcher = [func, pf, G](double* rhs, double c, ScratchPool&) {
        if (!pf->d_curr) throw std::runtime_error("pw GPU: source field not on device");
        hipLaunchKernelGGL(detail::pw1Kernel<Functor>, detail::pwGrid(G), dim3(64), 0, 0, rhs, pf->d_curr, func, c,
                           G);
        detail::checkLaunch("pw GPU kernel");
    };
    t.cpu_kernel = [func, pf, G](double* rhs, double c, ScratchPool&) {
        const double* s = pf->curr.data();
        for (int k = 0; k < G.nz; ++k)
            for (int j = 0; j < G.ny; ++j)
                for (int i = 0; i < G.nx; ++i) {
                    long long q = detail::pwIndex(G, i, j, k);
                    rhs[q] += c * func(s[q]);
                }
    };
    return t;
}

template <typename Functor>
Term pw(const ScalarField& f1, const ScalarField& f2, Functor func, double coeff) {
    detail::requireSameShape(f1, f2, "pw(f1, f2)");
    Term t;
    t.type = TermType::POINTWISE;
    t.field = &f1;
    t.coeff = coeff;
    t.sym.kind = TermSym::PW2;
    t.sym.f[0] = &f1;
    t.sym.f[1] = &f2;
    t.sym.abb = detail::probeAbb(func);
    const ScalarField *p1 = &f1, *p2 = &f2;
    const detail::PwGeom G = detail::pwGeom(f1);
    t.gpu_launcher = [func, p1, p2, G](double* rhs, double c, ScratchPool&) {
        if (!p1->d_curr || !p2->d_curr)
            throw std::runtime_error("pw(f1,f2) GPU: a field not on device. Call allocDevice() and uploadToDevice() first.");
        hipLaunchKernelGGL(detail::pw2Kernel<Functor>, detail::pwGrid(G), dim3(64), 0, 0, rhs, p1->d_curr,
                           p2->d_curr, func, c, G);
        detail::checkLaunch("pw2 GPU kernel");
    };
    t.cpu_kernel = [func, p1, p2, G](double* rhs, double c, ScratchPool&) {
        const double *a = p1->curr.data(), *b = p2->curr.data();
        for (int k = 0; k < G.nz; ++k)
            for (int j = 0; j < G.ny; ++j)
                for (int i = 0; i < G.nx; ++i) {
                    long long q = detail::pwIndex(G, i, j, k);
                    rhs[q] += c * func(a[q], b[q]);
                }
    };
    return t;
}

template <typename Functor>
Term pw(const ScalarField& f1, const ScalarField& f2, const ScalarField& f3, Functor func, double coeff) {
    detail::requireSameShape(f1, f2, "pw(f1,f2,f3)");
    detail::requireSameShape(f1, f3, "pw(f1,f2,f3)");
    Term t;
    t.type = TermType::POINTWISE;
    t.field = &f1;
    t.coeff = coeff;
    t.sym.kind = TermSym::PW3;
    t.sym.f[0] = &f1;
    t.sym.f[1] = &f2;
    t.sym.f[2] = &f3;
    const ScalarField *p1 = &f1, *p2 = &f2, *p3 = &f3;
    const detail::PwGeom G = detail::pwGeom(f1);
    t.gpu_launcher = [func, p1, p2, p3, G](double* rhs, double c, ScratchPool&) {
        if (!p1->d_curr || !p2->d_curr || !p3->d_curr)
            throw std::runtime_error("pw(f1,f2,f3) GPU: a field not on device.");
        hipLaunchKernelGGL(detail::pw3Kernel<Functor>, detail::pwGrid(G), dim3(64), 0, 0, rhs, p1->d_curr,
                           p2->d_curr, p3->d_curr, func, c, G);
        detail::checkLaunch("pw3 GPU kernel");
    };
    t.cpu_kernel = [func, p1, p2, p3, G](double* rhs, double c, ScratchPool&) {
        const double *a = p1->curr.data(), *b = p2->curr.data(), *d = p3->curr.data();
        for (int k = 0; k < G.nz; ++k)
            for (int j = 0; j < G.ny; ++j)
                for (int i = 0; i < G.nx; ++i) {
                    long long q = detail::pwIndex(G, i, j, k);
                    rhs[q] += c * func(a[q], b[q], d[q]);
                }
    };
    return t;
}

template <typename Functor>
VectorRHSExpr pw(const VectorField& vf, Functor func, double coeff) {
    VectorRHSExpr e(vf.nComponents());
    for (int c = 0; c < vf.nComponents(); ++c) e[c] = RHSExpr(pw(vf[c], func, coeff));
    return e;
}

template <typename Functor>
VectorRHSExpr pw(const VectorField& vf, const ScalarField& sf, Functor func, double coeff) {
    VectorRHSExpr e(vf.nComponents());
    for (int c = 0; c < vf.nComponents(); ++c) e[c] = RHSExpr(pw(vf[c], sf, func, coeff));
    return e;
}

template <typename Functor>
VectorRHSExpr pw(const VectorField& vf1, const VectorField& vf2, Functor func, double coeff) {
    if (vf1.nComponents() != vf2.nComponents())
        throw std::invalid_argument("pw(vf1, vf2): VectorFields must have the same number of components");
    VectorRHSExpr e(vf1.nComponents());
    for (int c = 0; c < vf1.nComponents(); ++c) e[c] = RHSExpr(pw(vf1[c], vf2[c], func, coeff));
    return e;
}

}  // namespace PhiX

#include "equation/FieldOps.inl"
