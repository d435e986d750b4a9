// equation.hip -- DSL factories, ScratchPool, Equation and VectorEquation of
// the drop-in C++ API (reference: src/equation/Equation.cu:130-745,
// src/equation/VectorEquation.cu:8-61).
//
// Device launchers call the FP64 term kernels of the C ABI
// (phx_lap_accumulate, phx_grad_accumulate, phx_mul_accumulate) on the null
// stream; host twins loop over the physical cells in the same order. Both keep
// the reference's operation order, so the generic path matches the FP64 oracle.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <algorithm>
#include <stdexcept>
#include <string>

#include "../../../include/phix_amd.h"
#include "boundary/BoundaryCondition.h"
#include "equation/Equation.h"
#include "equation/VectorEquation.h"

namespace PhiX {
namespace detail {

phx_grid gridOf(const ScalarField& f);
void throwPhx(int rc, const std::string& who);

void checkLaunch(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + " error: " + hipGetErrorString(e));
}

static void hipCheck(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

struct HostGeom {
    int nx, ny, nz, g, dim;
    long long sx, sxy;
    double idx2, idy2, idz2;
};

static HostGeom hostGeom(const ScalarField& f) {
    HostGeom h;
    h.nx = f.mesh.n[0];
    h.ny = f.mesh.n[1];
    h.nz = f.mesh.n[2];
    h.g = f.ghost;
    h.dim = f.mesh.dim;
    h.sx = f.storedDims[0];
    h.sxy = h.sx * f.storedDims[1];
    h.idx2 = 1.0 / (f.mesh.d[0] * f.mesh.d[0]);
    h.idy2 = h.dim >= 2 ? 1.0 / (f.mesh.d[1] * f.mesh.d[1]) : 0.0;
    h.idz2 = h.dim >= 3 ? 1.0 / (f.mesh.d[2] * f.mesh.d[2]) : 0.0;
    return h;
}

template <typename Fn>
static void forPhys(const HostGeom& h, Fn fn) {
    for (int k = 0; k < h.nz; ++k)
        for (int j = 0; j < h.ny; ++j)
            for (int i = 0; i < h.nx; ++i) fn((i + h.g) + h.sx * (j + h.g) + h.sxy * (k + h.g));
}

static void lapCPU(const HostGeom& h, double* rhs, const double* s, double c) {
    forPhys(h, [&](long long q) {
        double val = (s[q + 1] - 2.0 * s[q] + s[q - 1]) * h.idx2;
        if (h.dim >= 2) val += (s[q + h.sx] - 2.0 * s[q] + s[q - h.sx]) * h.idy2;
        if (h.dim >= 3) val += (s[q + h.sxy] - 2.0 * s[q] + s[q - h.sxy]) * h.idz2;
        rhs[q] += c * val;
    });
}

static void gradCPU(const HostGeom& h, double* rhs, const double* s, int axis, double inv2d, double c) {
    const long long st = axis == 0 ? 1 : (axis == 1 ? h.sx : h.sxy);
    forPhys(h, [&](long long q) { rhs[q] += c * (s[q + st] - s[q - st]) * inv2d; });
}

void mulAccumulateGPU(double* rhs, const double* s1, const double* s2, double c, const ScalarField& layout) {
    phx_grid g = gridOf(layout);
    throwPhx(phx_mul_accumulate(&g, rhs, s1, s2, c, nullptr), "mul_accumulate GPU");
}

void mulAccumulateCPU(double* rhs, const double* s1, const double* s2, double c, const ScalarField& layout) {
    forPhys(hostGeom(layout), [&](long long q) { rhs[q] += c * s1[q] * s2[q]; });
}

const ScalarField* repField(const Term& t) {
    if (!t.field) throw std::runtime_error("Composite Term: no representative source field captured");
    return t.field;
}
const ScalarField* repField(const RHSExpr& e) {
    for (const auto& t : e.terms)
        if (t.field) return t.field;
    throw std::runtime_error("Composite RHSExpr: no representative source field captured");
}

void materialiseGPU(const RHSExpr& e, double* buf, std::size_t n, ScratchPool& pool) {
    hipCheck(hipMemset(buf, 0, n * sizeof(double)), "materialise GPU memset");
    for (const auto& t : e.terms) {
        if (!t.gpu_launcher) throw std::runtime_error("Composite GPU: a Term has no GPU launcher");
        t.gpu_launcher(buf, t.coeff, pool);
    }
}
void materialiseGPU(const Term& t, double* buf, std::size_t n, ScratchPool& pool) {
    materialiseGPU(RHSExpr(t), buf, n, pool);
}
void materialiseCPU(const RHSExpr& e, double* buf, std::size_t n, ScratchPool& pool) {
    std::fill(buf, buf + n, 0.0);
    for (const auto& t : e.terms) {
        if (!t.cpu_kernel) throw std::runtime_error("Composite CPU: a Term has no CPU kernel");
        t.cpu_kernel(buf, t.coeff, pool);
    }
}
void materialiseCPU(const Term& t, double* buf, std::size_t n, ScratchPool& pool) {
    materialiseCPU(RHSExpr(t), buf, n, pool);
}

// Composite stencil: materialise `child` into scratch, BCs on it, then op.
static Term stencilOnExpr(const RHSExpr& child, const ScalarField& layout, const std::vector<BoundaryCondition*>& bcs,
                          double coeff, int axis, bool isLap) {
    if (!isLap && (axis < 0 || axis >= layout.mesh.dim)) throw std::invalid_argument("grad(expr): axis out of range");
    Term out;
    out.type = TermType::COMPOSITE;
    out.coeff = coeff;
    out.axis = isLap ? 0 : axis;
    out.field = &layout;
    out.sym.kind = isLap ? TermSym::LAP_EXPR : TermSym::GRAD_EXPR;
    out.sym.child = std::make_shared<std::vector<Term>>(child.terms);
    out.sym.bcs = bcs;
    out.sym.f[0] = &layout;
    const ScalarField* pl = &layout;
    const double inv2d = isLap ? 0.0 : 0.5 / layout.mesh.d[axis];
    out.gpu_launcher = [child, pl, bcs, isLap, axis](double* rhs, double c, ScratchPool& pool) {
        double* s = pool.acquireDevice(pl->storedSize);
        materialiseGPU(child, s, pl->storedSize, pool);
        ScalarField shell = ScalarField::makeShell(pl->mesh, pl->ghost, s);
        for (auto* bc : bcs) bc->applyOnGPU(shell);
        phx_grid g = gridOf(*pl);
        if (isLap) throwPhx(phx_lap_accumulate(&g, rhs, s, c, nullptr), "lap(expr) GPU");
        else throwPhx(phx_grad_accumulate(&g, rhs, s, axis, c, nullptr), "grad(expr) GPU");
    };
    out.cpu_kernel = [child, pl, bcs, isLap, axis, inv2d](double* rhs, double c, ScratchPool& pool) {
        double* s = pool.acquireHost(pl->storedSize);
        materialiseCPU(child, s, pl->storedSize, pool);
        ScalarField tmp(pl->mesh, "shell_cpu", pl->ghost);
        std::copy(s, s + pl->storedSize, tmp.curr.begin());
        for (auto* bc : bcs) bc->applyOnCPU(tmp);
        std::copy(tmp.curr.begin(), tmp.curr.end(), s);
        HostGeom h = hostGeom(*pl);
        if (isLap) lapCPU(h, rhs, s, c);
        else gradCPU(h, rhs, s, axis, inv2d, c);
    };
    return out;
}

}  // namespace detail

// ----------------------------------------------------------- ScratchPool ----
ScratchPool::~ScratchPool() {
    for (double* p : dev_bufs_)
        if (p) (void)hipFree(p);
}

double* ScratchPool::acquireDevice(std::size_t size) {
    if (next_dev_ < dev_bufs_.size()) {
        if (dev_sizes_[next_dev_] < size) {
            (void)hipFree(dev_bufs_[next_dev_]);
            dev_bufs_[next_dev_] = nullptr;
            detail::hipCheck(hipMalloc((void**)&dev_bufs_[next_dev_], size * sizeof(double)), "ScratchPool hipMalloc");
            dev_sizes_[next_dev_] = size;
        }
        return dev_bufs_[next_dev_++];
    }
    double* p = nullptr;
    detail::hipCheck(hipMalloc((void**)&p, size * sizeof(double)), "ScratchPool hipMalloc");
    dev_bufs_.push_back(p);
    dev_sizes_.push_back(size);
    ++next_dev_;
    return p;
}

double* ScratchPool::acquireHost(std::size_t size) {
    if (next_host_ < host_bufs_.size()) {
        if (host_bufs_[next_host_].size() < size) host_bufs_[next_host_].assign(size, 0.0);
        return host_bufs_[next_host_++].data();
    }
    host_bufs_.emplace_back(size, 0.0);
    ++next_host_;
    return host_bufs_.back().data();
}

// ------------------------------------------------------------ factories ----
Term lap(const ScalarField& f, double coeff) {
    Term t;
    t.type = TermType::LAPLACIAN;
    t.field = &f;
    t.coeff = coeff;
    t.sym.kind = TermSym::LAP;
    t.sym.f[0] = &f;
    const ScalarField* pf = &f;
    t.gpu_launcher = [pf](double* rhs, double c, ScratchPool&) {
        if (!pf->d_curr) throw std::runtime_error("lap GPU: source field not on device");
        phx_grid g = detail::gridOf(*pf);
        detail::throwPhx(phx_lap_accumulate(&g, rhs, pf->d_curr, c, nullptr), "lap GPU");
    };
    t.cpu_kernel = [pf](double* rhs, double c, ScratchPool&) {
        detail::lapCPU(detail::hostGeom(*pf), rhs, pf->curr.data(), c);
    };
    return t;
}

Term grad(const ScalarField& f, int axis, double coeff) {
    if (axis < 0 || axis >= f.mesh.dim) throw std::invalid_argument("grad: axis out of range for this mesh dimension");
    Term t;
    t.type = TermType::GRADIENT;
    t.field = &f;
    t.coeff = coeff;
    t.axis = axis;
    t.sym.kind = TermSym::GRAD;
    t.sym.f[0] = &f;
    const ScalarField* pf = &f;
    const double inv2d = 0.5 / f.mesh.d[axis];
    t.gpu_launcher = [pf, axis](double* rhs, double c, ScratchPool&) {
        if (!pf->d_curr) throw std::runtime_error("grad GPU: source field not on device");
        phx_grid g = detail::gridOf(*pf);
        detail::throwPhx(phx_grad_accumulate(&g, rhs, pf->d_curr, axis, c, nullptr), "grad GPU");
    };
    t.cpu_kernel = [pf, axis, inv2d](double* rhs, double c, ScratchPool&) {
        detail::gradCPU(detail::hostGeom(*pf), rhs, pf->curr.data(), axis, inv2d, c);
    };
    return t;
}

Term lap(const Term& t, const std::vector<BoundaryCondition*>& bcs, double coeff) {
    return detail::stencilOnExpr(RHSExpr(t), *detail::repField(t), bcs, coeff, 0, true);
}
Term lap(const RHSExpr& e, const std::vector<BoundaryCondition*>& bcs, double coeff) {
    return detail::stencilOnExpr(e, *detail::repField(e), bcs, coeff, 0, true);
}
Term grad(const Term& t, int axis, const std::vector<BoundaryCondition*>& bcs, double coeff) {
    return detail::stencilOnExpr(RHSExpr(t), *detail::repField(t), bcs, coeff, axis, false);
}
Term grad(const RHSExpr& e, int axis, const std::vector<BoundaryCondition*>& bcs, double coeff) {
    return detail::stencilOnExpr(e, *detail::repField(e), bcs, coeff, axis, false);
}

VectorRHSExpr grad(const Term& t, const std::vector<BoundaryCondition*>& bcs, double coeff) {
    const int dim = detail::repField(t)->mesh.dim;
    VectorRHSExpr v(dim);
    for (int a = 0; a < dim; ++a) v[a] = RHSExpr(grad(t, a, bcs, coeff));
    return v;
}
VectorRHSExpr grad(const RHSExpr& e, const std::vector<BoundaryCondition*>& bcs, double coeff) {
    const int dim = detail::repField(e)->mesh.dim;
    VectorRHSExpr v(dim);
    for (int a = 0; a < dim; ++a) v[a] = RHSExpr(grad(e, a, bcs, coeff));
    return v;
}

VectorRHSExpr lap(const VectorField& vf, double coeff) {
    VectorRHSExpr v(vf.nComponents());
    for (int c = 0; c < vf.nComponents(); ++c) v[c] = RHSExpr(lap(vf[c], coeff));
    return v;
}
VectorRHSExpr grad(const ScalarField& f, double coeff) {
    VectorRHSExpr v(f.mesh.dim);
    for (int a = 0; a < f.mesh.dim; ++a) v[a] = RHSExpr(grad(f, a, coeff));
    return v;
}
RHSExpr div(const VectorField& vf, double coeff) {
    if (vf.nComponents() < vf.mesh.dim)
        throw std::invalid_argument("div: VectorField must have at least mesh.dim components");
    RHSExpr e;
    for (int a = 0; a < vf.mesh.dim; ++a) e += grad(vf[a], a, coeff);
    return e;
}
RHSExpr div(const VectorRHSExpr& v, const std::vector<BoundaryCondition*>& bcs, double coeff) {
    RHSExpr e;
    for (int a = 0; a < v.nComponents(); ++a) e += grad(v[a], a, bcs, coeff);
    return e;
}
VectorRHSExpr curl(const VectorField& vf, double coeff) {
    if (vf.mesh.dim != 3 || vf.nComponents() != 3)
        throw std::invalid_argument("curl: VectorField must be 3-component on a 3D mesh");
    VectorRHSExpr v(3);
    v[0] += grad(vf[2], 1, coeff);
    v[0] += grad(vf[1], 2, -coeff);
    v[1] += grad(vf[0], 2, coeff);
    v[1] += grad(vf[2], 0, -coeff);
    v[2] += grad(vf[1], 0, coeff);
    v[2] += grad(vf[0], 1, -coeff);
    return v;
}

// -------------------------------------------------------------- Equation ----
Equation::Equation(ScalarField& u, const std::string& nm) : name(nm), unknown(u) {}

void Equation::setRHS(const RHSExpr& e) { rhs_expr_ = e; }
void Equation::setRHS(const Term& t) { rhs_expr_ = RHSExpr(t); }

void Equation::computeRHS(ScalarField& rhs) const {
    enqueueRHS(rhs);
    detail::hipCheck(hipDeviceSynchronize(), "Equation::computeRHS");
}

void Equation::enqueueRHS(ScalarField& rhs) const {
    if (!rhs.deviceAllocated()) throw std::runtime_error("Equation::computeRHS: rhs device memory not allocated");
    if (rhs_expr_.terms.empty()) throw std::runtime_error("Equation::computeRHS: RHS not set (call setRHS first)");
    detail::hipCheck(hipMemsetAsync(rhs.d_curr, 0, rhs.storedSize * sizeof(double), nullptr),
                     "Equation::computeRHS memset");
    scratch_pool_.reset();
    for (const auto& t : rhs_expr_.terms) {
        if (!t.gpu_launcher) throw std::runtime_error("Equation::computeRHS: a Term has no GPU launcher.");
        t.gpu_launcher(rhs.d_curr, t.coeff, scratch_pool_);
    }
    // Launch errors surface here, faults at the caller's next wait.
    detail::hipCheck(hipGetLastError(), "Equation::computeRHS");
}

void Equation::computeRHSCPU(ScalarField& rhs) const {
    if (rhs_expr_.terms.empty()) throw std::runtime_error("Equation::computeRHSCPU: RHS not set");
    std::fill(rhs.curr.begin(), rhs.curr.end(), 0.0);
    scratch_pool_.reset();
    for (const auto& t : rhs_expr_.terms) {
        if (!t.cpu_kernel) throw std::runtime_error("Equation::computeRHSCPU: a Term has no CPU kernel.");
        t.cpu_kernel(rhs.curr.data(), t.coeff, scratch_pool_);
    }
}

// -------------------------------------------------------- VectorEquation ----
VectorEquation::VectorEquation(VectorField& u, const std::string& nm) : name(nm), unknown(u) {
    for (int c = 0; c < u.nComponents(); ++c)
        equations_.push_back(std::make_unique<Equation>(u[c], nm + "_c" + std::to_string(c)));
}

void VectorEquation::setRHS(const VectorRHSExpr& e) {
    const int n = (int)equations_.size();
    if (e.nComponents() != n)
        throw std::invalid_argument("VectorEquation::setRHS: VectorRHSExpr has " + std::to_string(e.nComponents()) +
                                    " components but unknown has " + std::to_string(n));
    for (int c = 0; c < n; ++c) equations_[c]->setRHS(e[c]);
}

void VectorEquation::computeRHS(VectorField& rhs) const {
    if (rhs.nComponents() != (int)equations_.size())
        throw std::invalid_argument("VectorEquation::computeRHS: rhs component count mismatch");
    for (std::size_t c = 0; c < equations_.size(); ++c) equations_[c]->computeRHS(rhs[(int)c]);
}

void VectorEquation::computeRHSCPU(VectorField& rhs) const {
    if (rhs.nComponents() != (int)equations_.size())
        throw std::invalid_argument("VectorEquation::computeRHSCPU: rhs component count mismatch");
    for (std::size_t c = 0; c < equations_.size(); ++c) equations_[c]->computeRHSCPU(rhs[(int)c]);
}

Equation& VectorEquation::componentEquation(int c) { return *equations_.at(c); }
const Equation& VectorEquation::componentEquation(int c) const { return *equations_.at(c); }

bool VectorEquation::hasRHS() const {
    if (equations_.empty()) return false;
    for (const auto& e : equations_)
        if (!e->hasRHS()) return false;
    return true;
}

}  // namespace PhiX
