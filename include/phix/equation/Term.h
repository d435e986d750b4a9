#pragma once
// Expression DSL for right-hand sides; drop-in for include/equation/Term.h:24-326
// of the reference. A Term is one additive contribution with a coefficient and
// two launchers (device and host); an RHSExpr is an ordered sum of Terms.
//
// Each Term also carries a TermSym: what the term *is* (lap of a field, a
// pointwise double-well of a field, a composite lap(expr, bcs), ...). The
// Solver's planner reads it to replace a whole RHS by one fused kernel; terms
// it cannot name are OPAQUE and run through their launchers exactly as in the
// reference.
//
// Needs hipcc (pw() instantiates device kernels for the user's functors).

#include <hip/hip_runtime.h>

#include <cstddef>
#include <functional>
#include <memory>
#include <vector>

#include "field/ScalarField.h"
#include "field/VectorField.h"

// pw() lambdas must be callable on host and device.
#define PHIX_FN [=] __host__ __device__

namespace PhiX {

class BoundaryCondition;
struct Term;

// Recyclable scratch buffers for composite terms (sizes in doubles). Reset by
// Equation::computeRHS before each evaluation; buffers are not zeroed.
class ScratchPool {
public:
    ScratchPool() = default;
    ~ScratchPool();
    ScratchPool(const ScratchPool&) = delete;
    ScratchPool& operator=(const ScratchPool&) = delete;

    double* acquireDevice(std::size_t size);
    double* acquireHost(std::size_t size);
    void reset() { next_dev_ = 0; next_host_ = 0; }

private:
    std::vector<double*> dev_bufs_;
    std::vector<std::size_t> dev_sizes_;
    std::size_t next_dev_ = 0;
    std::vector<std::vector<double>> host_bufs_;
    std::size_t next_host_ = 0;
};

// (rhs, effective coefficient, scratch pool). Launchers hold `const ScalarField*`
// and read d_curr / curr when invoked, so swapping a field's buffer pointer
// between calls redirects them.
using TermLauncher = std::function<void(double*, double, ScratchPool&)>;

enum class TermType { LAPLACIAN, GRADIENT, POINTWISE, COMPOSITE };

// What the planner knows about a term.
struct TermSym {
    enum Kind { OPAQUE = 0, LAP, GRAD, PW1, PW2, PW3, MUL, LAP_EXPR, GRAD_EXPR };
    int kind = OPAQUE;
    const ScalarField* f[3] = {nullptr, nullptr, nullptr};
    // PW1: F(c) reproduces 2*rho*(c-ca)*(c-cb)*(2c-ca-cb) bit for bit at the probe points
    bool doublewell = false;
    double dw[3] = {0.0, 0.0, 0.0};     // rho, ca, cb
    // PW1: F(e) == e - e*e*e bit for bit at the probe points
    bool ac_well = false;
    // PW2: F(a,b) == a*b*b bit for bit at the probe points
    bool abb = false;
    // LAP_EXPR / GRAD_EXPR: the materialised child and the BCs applied to it
    std::shared_ptr<std::vector<Term>> child;
    std::vector<BoundaryCondition*> bcs;
};

struct Term {
    TermType type = TermType::LAPLACIAN;
    double coeff = 1.0;
    const ScalarField* field = nullptr;
    int axis = 0;

    TermLauncher gpu_launcher;
    TermLauncher cpu_kernel;

    TermSym sym;

    Term operator*(double s) const { Term t = *this; t.coeff *= s; return t; }
    Term operator/(double s) const { return *this * (1.0 / s); }
    Term operator-() const { return *this * (-1.0); }
};

inline Term operator*(double s, const Term& t) { return t * s; }
inline Term operator/(const Term& t, double s) { return t * (1.0 / s); }

struct RHSExpr {
    std::vector<Term> terms;

    RHSExpr() = default;
    explicit RHSExpr(const Term& t) { terms.push_back(t); }

    RHSExpr& operator+=(const Term& t) { terms.push_back(t); return *this; }
    RHSExpr& operator-=(const Term& t) { terms.push_back(-t); return *this; }
    RHSExpr& operator+=(const RHSExpr& e) {
        terms.insert(terms.end(), e.terms.begin(), e.terms.end());
        return *this;
    }
    RHSExpr& operator-=(const RHSExpr& e) {
        for (const auto& t : e.terms) terms.push_back(-t);
        return *this;
    }
    RHSExpr operator+(const Term& t) const { RHSExpr r = *this; r += t; return r; }
    RHSExpr operator-(const Term& t) const { RHSExpr r = *this; r -= t; return r; }
    RHSExpr operator+(const RHSExpr& e) const { RHSExpr r = *this; r += e; return r; }
    RHSExpr operator-(const RHSExpr& e) const { RHSExpr r = *this; r -= e; return r; }
    RHSExpr operator*(double s) const {
        RHSExpr r;
        for (const auto& t : terms) r.terms.push_back(t * s);
        return r;
    }
};

inline RHSExpr operator+(const Term& a, const Term& b) { RHSExpr e(a); e += b; return e; }
inline RHSExpr operator-(const Term& a, const Term& b) { RHSExpr e(a); e -= b; return e; }
inline RHSExpr operator+(const Term& t, const RHSExpr& e) { RHSExpr r(t); r += e; return r; }
inline RHSExpr operator-(const Term& t, const RHSExpr& e) { RHSExpr r(t); r -= e; return r; }
inline RHSExpr operator*(double s, const RHSExpr& e) { return e * s; }

// coeff * lap(f): second-order central differences summed over active axes.
Term lap(const ScalarField& f, double coeff = 1.0);
// coeff * d f / d x_axis: second-order central difference.
Term grad(const ScalarField& f, int axis, double coeff = 1.0);

// Operators on composite expressions: the expression is materialised into a
// scratch field, `bcs` fill its ghosts, then the stencil runs on it.
Term lap(const Term& t, const std::vector<BoundaryCondition*>& bcs, double coeff = 1.0);
Term lap(const RHSExpr& expr, const std::vector<BoundaryCondition*>& bcs, double coeff = 1.0);
Term grad(const Term& t, int axis, const std::vector<BoundaryCondition*>& bcs, double coeff = 1.0);
Term grad(const RHSExpr& expr, int axis, const std::vector<BoundaryCondition*>& bcs, double coeff = 1.0);

struct VectorRHSExpr;
VectorRHSExpr grad(const Term& t, const std::vector<BoundaryCondition*>& bcs, double coeff = 1.0);
VectorRHSExpr grad(const RHSExpr& expr, const std::vector<BoundaryCondition*>& bcs, double coeff = 1.0);

// Pointwise user transforms: rhs += coeff * func(f[idx]) (1, 2 or 3 fields).
template <typename Functor>
Term pw(const ScalarField& f, Functor func, double coeff = 1.0);
template <typename Functor>
Term pw(const ScalarField& f1, const ScalarField& f2, Functor func, double coeff = 1.0);
template <typename Functor>
Term pw(const ScalarField& f1, const ScalarField& f2, const ScalarField& f3, Functor func, double coeff = 1.0);

struct VectorRHSExpr {
    std::vector<RHSExpr> components;

    VectorRHSExpr() = default;
    explicit VectorRHSExpr(int n) : components(n) {}

    int nComponents() const { return static_cast<int>(components.size()); }
    RHSExpr& operator[](int c) { return components[c]; }
    const RHSExpr& operator[](int c) const { return components[c]; }

    VectorRHSExpr& operator+=(const VectorRHSExpr& o) {
        for (int c = 0; c < nComponents(); ++c) components[c] += o.components[c];
        return *this;
    }
    VectorRHSExpr& operator-=(const VectorRHSExpr& o) {
        for (int c = 0; c < nComponents(); ++c) components[c] -= o.components[c];
        return *this;
    }
    VectorRHSExpr operator+(const VectorRHSExpr& o) const { VectorRHSExpr r = *this; r += o; return r; }
    VectorRHSExpr operator-(const VectorRHSExpr& o) const { VectorRHSExpr r = *this; r -= o; return r; }
    VectorRHSExpr operator*(double s) const {
        VectorRHSExpr r(nComponents());
        for (int c = 0; c < nComponents(); ++c) r[c] = components[c] * s;
        return r;
    }
    VectorRHSExpr operator-() const { return (*this) * (-1.0); }
};

inline VectorRHSExpr operator*(double s, const VectorRHSExpr& e) { return e * s; }

VectorRHSExpr lap(const VectorField& vf, double coeff = 1.0);
VectorRHSExpr grad(const ScalarField& f, double coeff = 1.0);
RHSExpr div(const VectorField& vf, double coeff = 1.0);
RHSExpr div(const VectorRHSExpr& v, const std::vector<BoundaryCondition*>& bcs, double coeff = 1.0);
VectorRHSExpr curl(const VectorField& vf, double coeff = 1.0);

template <typename Functor>
VectorRHSExpr pw(const VectorField& vf, Functor func, double coeff = 1.0);
template <typename Functor>
VectorRHSExpr pw(const VectorField& vf, const ScalarField& sf, Functor func, double coeff = 1.0);
template <typename Functor>
VectorRHSExpr pw(const VectorField& vf1, const VectorField& vf2, Functor func, double coeff = 1.0);

}  // namespace PhiX

#include "equation/TermPW.inl"
