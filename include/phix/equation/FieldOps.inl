// FieldOps.inl -- arithmetic between fields, scalars, Terms and RHSExprs
// (included by TermPW.inl). Same operator set and meaning as the reference
// (include/equation/FieldOps.inl:55-289):
//   field (+,-,*) field, field (+,-,*) scalar, -field          -> pw() Terms
//   Term/RHSExpr * field, Term/RHSExpr * Term/RHSExpr           -> COMPOSITE Terms
//     that materialise each expression operand into a scratch buffer and then
//     accumulate coeff*s1*s2
//   Term/field/RHSExpr * VectorRHSExpr                         -> per component
#pragma once

#include <stdexcept>

namespace PhiX {

namespace detail {
void mulAccumulateGPU(double* d_rhs, const double* d_s1, const double* d_s2, double coeff, const ScalarField& layout);
void mulAccumulateCPU(double* rhs, const double* s1, const double* s2, double coeff, const ScalarField& layout);
void materialiseGPU(const RHSExpr&, double*, std::size_t, ScratchPool&);
void materialiseGPU(const Term&, double*, std::size_t, ScratchPool&);
void materialiseCPU(const RHSExpr&, double*, std::size_t, ScratchPool&);
void materialiseCPU(const Term&, double*, std::size_t, ScratchPool&);
const ScalarField* repField(const Term&);
const ScalarField* repField(const RHSExpr&);

template <typename Lhs>
Term termTimesField(Lhs lhs, const ScalarField& field, const ScalarField& layout) {
    Term out;
    out.type = TermType::COMPOSITE;
    out.field = &layout;
    const ScalarField* pf = &field;
    const ScalarField* pl = &layout;
    out.gpu_launcher = [lhs, pf, pl](double* rhs, double c, ScratchPool& pool) {
        double* s = pool.acquireDevice(pl->storedSize);
        materialiseGPU(lhs, s, pl->storedSize, pool);
        if (!pf->d_curr) throw std::runtime_error("Term*Field GPU: field not on device");
        mulAccumulateGPU(rhs, s, pf->d_curr, c, *pl);
    };
    out.cpu_kernel = [lhs, pf, pl](double* rhs, double c, ScratchPool& pool) {
        double* s = pool.acquireHost(pl->storedSize);
        materialiseCPU(lhs, s, pl->storedSize, pool);
        mulAccumulateCPU(rhs, s, pf->curr.data(), c, *pl);
    };
    return out;
}

template <typename Lhs, typename Rhs>
Term termTimesTerm(Lhs lhs, Rhs r, const ScalarField& layout) {
    Term out;
    out.type = TermType::COMPOSITE;
    out.field = &layout;
    const ScalarField* pl = &layout;
    out.gpu_launcher = [lhs, r, pl](double* rhs, double c, ScratchPool& pool) {
        double* s1 = pool.acquireDevice(pl->storedSize);
        materialiseGPU(lhs, s1, pl->storedSize, pool);
        double* s2 = pool.acquireDevice(pl->storedSize);
        materialiseGPU(r, s2, pl->storedSize, pool);
        mulAccumulateGPU(rhs, s1, s2, c, *pl);
    };
    out.cpu_kernel = [lhs, r, pl](double* rhs, double c, ScratchPool& pool) {
        double* s1 = pool.acquireHost(pl->storedSize);
        materialiseCPU(lhs, s1, pl->storedSize, pool);
        double* s2 = pool.acquireHost(pl->storedSize);
        materialiseCPU(r, s2, pl->storedSize, pool);
        mulAccumulateCPU(rhs, s1, s2, c, *pl);
    };
    return out;
}
}  // namespace detail

inline Term operator*(const ScalarField& f1, const ScalarField& f2) {
    return pw(f1, f2, [] __host__ __device__(double a, double b) { return a * b; });
}
inline Term operator+(const ScalarField& f1, const ScalarField& f2) {
    return pw(f1, f2, [] __host__ __device__(double a, double b) { return a + b; });
}
inline Term operator-(const ScalarField& f1, const ScalarField& f2) {
    return pw(f1, f2, [] __host__ __device__(double a, double b) { return a - b; });
}
// s * f scales through the coefficient of an identity pw (unlike s * Term,
// which scales an existing Term's coefficient).
inline Term operator*(double s, const ScalarField& f) {
    return pw(f, [] __host__ __device__(double v) { return v; }, s);
}
inline Term operator*(const ScalarField& f, double s) {
    return pw(f, [] __host__ __device__(double v) { return v; }, s);
}
inline Term operator+(const ScalarField& f, double s) {
    return pw(f, [s] __host__ __device__(double v) { return v + s; });
}
inline Term operator+(double s, const ScalarField& f) {
    return pw(f, [s] __host__ __device__(double v) { return v + s; });
}
inline Term operator-(const ScalarField& f, double s) {
    return pw(f, [s] __host__ __device__(double v) { return v - s; });
}
inline Term operator-(double s, const ScalarField& f) {
    return pw(f, [s] __host__ __device__(double v) { return s - v; });
}
inline Term operator-(const ScalarField& f) {
    return pw(f, [] __host__ __device__(double v) { return -v; });
}

inline Term operator*(const Term& t, const ScalarField& f) { return detail::termTimesField<Term>(t, f, *detail::repField(t)); }
inline Term operator*(const ScalarField& f, const Term& t) { return t * f; }
inline Term operator*(const Term& a, const Term& b) {
    return detail::termTimesTerm<Term, Term>(a, b, *detail::repField(a));
}
inline Term operator*(const RHSExpr& e, const ScalarField& f) {
    return detail::termTimesField<RHSExpr>(e, f, *detail::repField(e));
}
inline Term operator*(const ScalarField& f, const RHSExpr& e) { return e * f; }
inline Term operator*(const RHSExpr& a, const Term& b) {
    return detail::termTimesTerm<RHSExpr, Term>(a, b, *detail::repField(a));
}
inline Term operator*(const Term& a, const RHSExpr& b) {
    return detail::termTimesTerm<Term, RHSExpr>(a, b, *detail::repField(a));
}
inline Term operator*(const RHSExpr& a, const RHSExpr& b) {
    return detail::termTimesTerm<RHSExpr, RHSExpr>(a, b, *detail::repField(a));
}

inline VectorRHSExpr operator*(const Term& t, const VectorRHSExpr& v) {
    VectorRHSExpr out(v.nComponents());
    for (int c = 0; c < v.nComponents(); ++c) out[c] = RHSExpr(t * RHSExpr(v[c]));
    return out;
}
inline VectorRHSExpr operator*(const VectorRHSExpr& v, const Term& t) { return t * v; }
inline VectorRHSExpr operator*(const ScalarField& f, const VectorRHSExpr& v) {
    VectorRHSExpr out(v.nComponents());
    for (int c = 0; c < v.nComponents(); ++c) out[c] = RHSExpr(f * RHSExpr(v[c]));
    return out;
}
inline VectorRHSExpr operator*(const VectorRHSExpr& v, const ScalarField& f) { return f * v; }
inline VectorRHSExpr operator*(const RHSExpr& e, const VectorRHSExpr& v) {
    VectorRHSExpr out(v.nComponents());
    for (int c = 0; c < v.nComponents(); ++c) out[c] = RHSExpr(e * RHSExpr(v[c]));
    return out;
}
inline VectorRHSExpr operator*(const VectorRHSExpr& v, const RHSExpr& e) { return e * v; }

}  // namespace PhiX
