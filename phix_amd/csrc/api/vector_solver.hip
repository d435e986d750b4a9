// vector_solver.hip -- VectorSolver of the drop-in C++ API
// (reference: src/solver/VectorSolver.cu). Generic plan: the reference
// algorithm -- all component pointers swapped to the stage state together,
// BCs on every stage state (VectorSolver.cu:149-228). Fused plan: the coupled
// Allen-Cahn form named in solver/VectorSolver.h, one kernel per RK4 stage
// for all components (phx_ac_rk4_steps).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <stdexcept>
#include <string>

#include "../../../include/phix_amd.h"
#include "solver/VectorSolver.h"

namespace PhiX {
namespace detail {
phx_grid gridOf(const ScalarField& f);
void throwPhx(int rc, const std::string& who);
bool bcTable(const std::vector<BoundaryCondition*>& bcs, const Mesh& m, phx_grid& g);

static void hipOk2(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// Component i: [lap(eta_i) cl, pw(eta_i, e-e^3) cp, pw(eta_i, eta_j, a*b*b) cg for j != i ascending]
static bool matchAC(const VectorEquation& eq, double& cl, double& cp, double& cg) {
    const VectorField& v = eq.unknown;
    const int N = v.nComponents();
    if (N < 1 || N > 8) return false;
    for (int i = 0; i < N; ++i) {
        const auto& t = eq.componentEquation(i).rhsExpr().terms;
        if ((int)t.size() != 1 + 1 + (N - 1)) return false;
        if (t[0].sym.kind != TermSym::LAP || t[0].sym.f[0] != &v[i]) return false;
        if (t[1].sym.kind != TermSym::PW1 || !t[1].sym.ac_well || t[1].sym.f[0] != &v[i]) return false;
        if (i == 0) {
            cl = t[0].coeff;
            cp = t[1].coeff;
        } else if (t[0].coeff != cl || t[1].coeff != cp) {
            return false;
        }
        int q = 2;
        for (int j = 0; j < N; ++j) {
            if (j == i) continue;
            const Term& x = t[q++];
            if (x.sym.kind != TermSym::PW2 || !x.sym.abb || x.sym.f[0] != &v[i] || x.sym.f[1] != &v[j]) return false;
            if (i == 0 && q == 3) cg = x.coeff;
            else if (x.coeff != cg) return false;
        }
    }
    return true;
}
}  // namespace detail

struct VectorSolver::Impl {
    bool fused = false;
    phx_grid grid{};
    phx_ac_params ac{};
    void* work = nullptr;
    std::vector<float*> s32;
    std::unique_ptr<VectorField> rhs, k1, k2, k3, k4, tmp;
    ~Impl() {
        if (work) (void)hipFree(work);
        for (float* p : s32)
            if (p) (void)hipFree(p);
    }
};

VectorSolver::VectorSolver(VectorEquation& eq, std::vector<BoundaryCondition*> bcs, double dt_, TimeScheme sch)
    : dt(dt_), scheme(sch), equation_(eq), bcs_(std::move(bcs)), use_rk4_(sch == TimeScheme::RK4), impl_(new Impl) {
    if (!equation_.unknown.deviceAllocated())
        throw std::runtime_error("VectorSolver: unknown must have device memory allocated before constructing the "
                                 "VectorSolver.");
    plan_name_ = "generic";
    Impl& I = *impl_;
    const VectorField& v = equation_.unknown;
    I.grid = detail::gridOf(v[0]);
    double cl = 0, cp = 0, cg = 0;
    if (use_rk4_ && v.mesh.dim <= 2 && v.ghost >= 1 && detail::bcTable(bcs_, v.mesh, I.grid) &&
        detail::matchAC(equation_, cl, cp, cg) && cp != 0.0) {
        // phx_ac_rk4_steps forms L*kappa, L and -2*L*gamma; accept only if they
        // reproduce the equation's coefficients exactly.
        double L = cp, kappa = cl / cp, gamma = cg / (-2.0 * cp);
        if (L * kappa == cl && -2.0 * L * gamma == cg) {
            I.fused = true;
            I.ac = phx_ac_params{L, kappa, gamma, dt};
            plan_name_ = "fused AC RK4 (phx_ac_rk4_steps)";
        }
    }
}

VectorSolver::~VectorSolver() = default;

void VectorSolver::setBulkPrecision(Precision p) {
    if (p == Precision::FP32 && !impl_->fused)
        throw std::invalid_argument("VectorSolver::setBulkPrecision: FP32 bulk needs the fused plan");
    precision_ = p;
}

static std::unique_ptr<VectorField> vscratch(const VectorField& ref, const std::string& tag) {
    auto f = std::make_unique<VectorField>(ref.mesh, tag, ref.nComponents(), ref.ghost);
    f->allocDevice();
    return f;
}

void VectorSolver::advanceN(int n) {
    Impl& I = *impl_;
    VectorField& phi = equation_.unknown;
    const int N = phi.nComponents();
    if (n <= 0) return;
    if (I.fused) {
        I.ac.dt = dt;
        const bool f32 = precision_ == Precision::FP32;
        const std::size_t S = phi[0].storedSize;
        std::vector<void*> ptrs(N);
        if (f32) {
            if (I.s32.empty()) {
                I.s32.assign(N, nullptr);
                for (int c = 0; c < N; ++c) detail::hipOk2(hipMalloc((void**)&I.s32[c], S * sizeof(float)), "FP32 state");
            }
            for (int c = 0; c < N; ++c) {
                detail::throwPhx(phx_convert_f64_to_f32(I.s32[c], phi[c].d_curr, S, nullptr), "FP32 in");
                ptrs[c] = I.s32[c];
            }
        } else {
            for (int c = 0; c < N; ++c) ptrs[c] = phi[c].d_curr;
        }
        if (!I.work)
            detail::hipOk2(hipMalloc(&I.work, phx_ac_rk4_workspace_bytes(&I.grid, PHX_F64, N)), "AC workspace");
        detail::throwPhx(phx_ac_rk4_steps(&I.grid, f32 ? PHX_F32 : PHX_F64, N, ptrs.data(), I.work, &I.ac, n, nullptr),
                         "VectorSolver::advance");
        if (f32)
            for (int c = 0; c < N; ++c)
                detail::throwPhx(phx_convert_f32_to_f64(phi[c].d_curr, I.s32[c], S, nullptr), "FP32 out");
        detail::hipOk2(hipDeviceSynchronize(), "VectorSolver::advance");
        phi.advanceTimeLevelGPU();
        return;
    }
    if (!I.rhs) {
        I.rhs = vscratch(phi, "_rhs");
        if (use_rk4_) {
            I.k1 = vscratch(phi, "_k1");
            I.k2 = vscratch(phi, "_k2");
            I.k3 = vscratch(phi, "_k3");
            I.k4 = vscratch(phi, "_k4");
            I.tmp = vscratch(phi, "_phi_tmp");
        }
    }
    for (int it = 0; it < n; ++it) {
        for (auto* bc : bcs_) bc->applyOnGPU(phi);
        if (!use_rk4_) {
            equation_.computeRHS(*I.rhs);
            for (int c = 0; c < N; ++c)
                detail::throwPhx(phx_axpy_f64(phi[c].d_curr, (*I.rhs)[c].d_curr, dt, phi[c].storedSize, nullptr),
                                 "Euler update");
        } else {
            VectorField* k[4] = {I.k1.get(), I.k2.get(), I.k3.get(), I.k4.get()};
            const double shift[3] = {dt * 0.5, dt * 0.5, dt};
            equation_.computeRHS(*k[0]);
            for (int s = 1; s < 4; ++s) {
                std::vector<double*> saved(N);
                for (int c = 0; c < N; ++c)
                    detail::throwPhx(phx_rk4_tmp_f64((*I.tmp)[c].d_curr, phi[c].d_curr, (*k[s - 1])[c].d_curr,
                                                     shift[s - 1], phi[c].storedSize, nullptr),
                                     "RK4 tmp");
                for (int c = 0; c < N; ++c) {
                    saved[c] = phi[c].d_curr;
                    phi[c].d_curr = (*I.tmp)[c].d_curr;
                }
                try {
                    for (auto* bc : bcs_) bc->applyOnGPU(phi);
                    equation_.computeRHS(*k[s]);
                } catch (...) {
                    for (int c = 0; c < N; ++c) phi[c].d_curr = saved[c];
                    throw;
                }
                for (int c = 0; c < N; ++c) phi[c].d_curr = saved[c];
            }
            for (int c = 0; c < N; ++c)
                detail::throwPhx(phx_rk4_update_f64(phi[c].d_curr, (*k[0])[c].d_curr, (*k[1])[c].d_curr,
                                                    (*k[2])[c].d_curr, (*k[3])[c].d_curr, dt / 6.0, phi[c].storedSize,
                                                    nullptr),
                                 "RK4 update");
        }
        detail::hipOk2(hipDeviceSynchronize(), "VectorSolver::advance");
        phi.advanceTimeLevelGPU();
    }
}

void VectorSolver::advance() {
    advanceN(1);
    ++step;
    time += dt;
}

void VectorSolver::run(int nSteps, int callbackEvery, std::function<void(const VectorSolver&)> callback) {
    int done = 0;
    while (done < nSteps) {
        int chunk = nSteps - done;
        if (callback && callbackEvery > 0) chunk = std::min(chunk, callbackEvery - (step % callbackEvery));
        advanceN(chunk);
        for (int q = 0; q < chunk; ++q) {
            ++step;
            time += dt;
        }
        done += chunk;
        if (callback && callbackEvery > 0 && step % callbackEvery == 0) callback(*this);
    }
}

void VectorSolver::advanceCPU() {
    VectorField& phi = equation_.unknown;
    const int N = phi.nComponents();
    VectorField rhs(phi.mesh, "_rhs_cpu", N, phi.ghost);
    if (!use_rk4_) {
        for (auto* bc : bcs_) bc->applyOnCPU(phi);
        equation_.computeRHSCPU(rhs);
        for (int c = 0; c < N; ++c)
            for (std::size_t i = 0; i < phi[c].curr.size(); ++i) phi[c].curr[i] += dt * rhs[c].curr[i];
    } else {
        std::vector<std::vector<double>> k[4];
        VectorField tmp(phi.mesh, "_tmp_cpu", N, phi.ghost);
        const double shift[3] = {dt * 0.5, dt * 0.5, dt};
        for (auto* bc : bcs_) bc->applyOnCPU(phi);
        equation_.computeRHSCPU(rhs);
        for (int c = 0; c < N; ++c) k[0].push_back(rhs[c].curr);
        for (int s = 1; s < 4; ++s) {
            for (int c = 0; c < N; ++c)
                for (std::size_t i = 0; i < phi[c].curr.size(); ++i)
                    tmp[c].curr[i] = phi[c].curr[i] + shift[s - 1] * k[s - 1][c][i];
            for (int c = 0; c < N; ++c) std::swap(phi[c].curr, tmp[c].curr);
            for (auto* bc : bcs_) bc->applyOnCPU(phi);
            equation_.computeRHSCPU(rhs);
            for (int c = 0; c < N; ++c) std::swap(phi[c].curr, tmp[c].curr);
            for (int c = 0; c < N; ++c) k[s].push_back(rhs[c].curr);
        }
        const double dt6 = dt / 6.0;
        for (int c = 0; c < N; ++c)
            for (std::size_t i = 0; i < phi[c].curr.size(); ++i)
                phi[c].curr[i] += dt6 * (k[0][c][i] + 2.0 * k[1][c][i] + 2.0 * k[2][c][i] + k[3][c][i]);
    }
    phi.advanceTimeLevelCPU();
    ++step;
    time += dt;
}

}  // namespace PhiX
