// solver.hip -- Solver of the drop-in C++ API (reference: src/solver/Solver.cu).
//
// The constructor plans the step (see solver/Solver.h). Fused plans call the
// C ABI with every BC turned into a per-face index map; the generic plan is
// the reference algorithm on the per-term kernels:
//   Euler      BCs(u); rhs = RHS(u); u += dt*rhs (stored array); prev <- curr
//   RK4        BCs on u and on every stage state, pointer swap of u.d_curr so
//              the term launchers read the stage (Solver.cu:192-259)
//   multi-step per step: BCs(source); STEADY u = RHS, TRANSIENT u += dt*RHS;
//              prev <- curr for TRANSIENT unknowns (Solver.cu:317-336)
// advance() and run() return with the device work complete, and prev == curr
// for every time-integrated unknown.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>

#include "../../../include/phix_amd.h"
#include "solver/Solver.h"

namespace PhiX {
namespace detail {
phx_grid gridOf(const ScalarField& f);
void throwPhx(int rc, const std::string& who);

static void hipOk(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// Per-face BC table of a BC list applied in order (last writer wins per
// face). Fails (returns false) on a kind the fused kernels do not know, a
// face left without a condition, a condition on an axis at or beyond the
// mesh dimension (applied object by object it fills ghost planes the table
// fill does not walk), or a list not in X, Y, Z order (the order fixes the
// corner ghosts, which the fused kernels fill X, Y, Z).
bool bcTable(const std::vector<BoundaryCondition*>& bcs, const Mesh& m, phx_grid& g) {
    for (int a = 0; a < 3; ++a)
        for (int s = 0; s < 2; ++s) {
            g.bc[a][s] = PHX_BC_NONE;
            g.bc_value[a][s] = 0.0;
        }
    int lastAxis = -1;
    for (auto* bc : bcs) {
        if (!bc) return false;
        const int a = (int)bc->axis;
        if (a < lastAxis || a >= m.dim) return false;
        lastAxis = a;
        const int k = bc->kind();
        if (k < PHX_BC_PERIODIC || k > PHX_BC_FIXED) return false;
        const bool lo = k == PHX_BC_PERIODIC || bc->side != Side::HIGH;
        const bool hi = k == PHX_BC_PERIODIC || bc->side != Side::LOW;
        if (lo) { g.bc[a][0] = k; g.bc_value[a][0] = bc->fixedValue(); }
        if (hi) { g.bc[a][1] = k; g.bc_value[a][1] = bc->fixedValue(); }
    }
    for (int a = 0; a < m.dim; ++a) {
        if (g.bc[a][0] == PHX_BC_NONE || g.bc[a][1] == PHX_BC_NONE) return false;
        if ((g.bc[a][0] == PHX_BC_PERIODIC) != (g.bc[a][1] == PHX_BC_PERIODIC)) return false;
        if (m.n[a] < 2) return false;
    }
    return true;
}

bool sameTable(const phx_grid& a, const phx_grid& b) {
    for (int x = 0; x < 3; ++x)
        for (int s = 0; s < 2; ++s)
            if (a.bc[x][s] != b.bc[x][s] || (a.bc[x][s] == PHX_BC_FIXED && a.bc_value[x][s] != b.bc_value[x][s]))
                return false;
    return true;
}

// mu = pw(c, doublewell) [coeff 1] + lap(c) [coeff b], either order.
bool matchMu(const std::vector<Term>& t, const ScalarField* c, double dw[3], double& b) {
    if (t.size() != 2) return false;
    const Term* pw = nullptr;
    const Term* lp = nullptr;
    for (const auto& x : t) {
        if (x.sym.kind == TermSym::PW1) pw = &x;
        else if (x.sym.kind == TermSym::LAP) lp = &x;
    }
    if (!pw || !lp) return false;
    if (pw->sym.f[0] != c || lp->sym.f[0] != c || !pw->sym.doublewell || pw->coeff != 1.0) return false;
    std::copy(pw->sym.dw, pw->sym.dw + 3, dw);
    b = lp->coeff;
    return true;
}

}  // namespace detail

using detail::hipOk;

struct Solver::Impl {
    enum Plan { GENERIC, CH_RK4, CH_EULER } plan = GENERIC;
    phx_grid grid{};
    phx_ch_params ch{};
    ScalarField* c = nullptr;
    ScalarField* mu = nullptr;
    // fused workspaces
    void* work = nullptr;
    float* c32 = nullptr;
    float* mu32 = nullptr;
    // generic scratch
    std::unique_ptr<ScalarField> rhs, k1, k2, k3, k4, tmp;
    std::vector<std::unique_ptr<ScalarField>> stepScratch;

    ~Impl() {
        if (work) (void)hipFree(work);
        if (c32) (void)hipFree(c32);
        if (mu32) (void)hipFree(mu32);
    }
};

static std::unique_ptr<ScalarField> scratchLike(const ScalarField& ref, const std::string& tag) {
    auto f = std::make_unique<ScalarField>(ref.mesh, tag, ref.ghost);
    f->allocDevice();
    return f;
}

Solver::Solver(Equation& equation, std::vector<BoundaryCondition*> bcs, double dt_, TimeScheme scheme_)
    : dt(dt_), scheme(scheme_), equation_(equation), bcs_(std::move(bcs)), use_rk4_(scheme_ == TimeScheme::RK4),
      impl_(new Impl) {
    if (!equation_.unknown.deviceAllocated())
        throw std::runtime_error(
            "Solver: equation.unknown must have device memory allocated before constructing the Solver.");
    Impl& I = *impl_;
    I.c = &equation_.unknown;
    I.grid = detail::gridOf(*I.c);
    plan_name_ = "generic";
    const auto& terms = equation_.rhsExpr().terms;
    // CH RK4: M*lap(pw(c,f') - kappa*lap(c), bcs), bcs equal to the solver's
    if (use_rk4_ && terms.size() == 1 && terms[0].sym.kind == TermSym::LAP_EXPR && terms[0].sym.child &&
        I.c->ghost >= 1 && detail::bcTable(bcs_, I.c->mesh, I.grid)) {
        phx_grid gmu = I.grid;
        double dw[3], b;
        if (detail::bcTable(terms[0].sym.bcs, I.c->mesh, gmu) && detail::sameTable(gmu, I.grid) &&
            detail::matchMu(*terms[0].sym.child, I.c, dw, b)) {
            I.plan = Impl::CH_RK4;
            I.ch = phx_ch_params{dw[0], dw[1], dw[2], -b, terms[0].coeff, dt};
            plan_name_ = "fused CH RK4 (phx_ch_rk4_steps)";
        }
    }
}

Solver::Solver(std::vector<SolverStep> steps, double dt_, TimeScheme scheme_)
    : dt(dt_), scheme(scheme_), equation_(*steps.at(0).equation), bcs_({}), use_rk4_(false), multiStep_(true),
      steps_(std::move(steps)), impl_(new Impl) {
    for (auto& s : steps_)
        if (!s.equation || !s.equation->unknown.deviceAllocated())
            throw std::runtime_error(
                "Solver (multi-step): all equation unknowns must have device memory allocated before constructing "
                "the Solver.");
    Impl& I = *impl_;
    plan_name_ = "generic multi-step";
    if (steps_.size() == 2 && steps_[0].type == EquationType::STEADY && steps_[1].type == EquationType::TRANSIENT) {
        ScalarField* c = steps_[0].sourceField;
        ScalarField* mu = &steps_[0].equation->unknown;
        const auto& t0 = steps_[0].equation->rhsExpr().terms;
        const auto& t1 = steps_[1].equation->rhsExpr().terms;
        double dw[3], b;
        phx_grid g0 = detail::gridOf(*c), g1 = g0;
        if (c && c != mu && steps_[1].sourceField == mu && &steps_[1].equation->unknown == c && c->ghost >= 1 &&
            mu->ghost == c->ghost && &mu->mesh == &c->mesh && detail::matchMu(t0, c, dw, b) && t1.size() == 1 &&
            t1[0].sym.kind == TermSym::LAP && t1[0].sym.f[0] == mu && detail::bcTable(steps_[0].bcs, c->mesh, g0) &&
            detail::bcTable(steps_[1].bcs, c->mesh, g1) && detail::sameTable(g0, g1)) {
            I.plan = Impl::CH_EULER;
            I.grid = g0;
            I.c = c;
            I.mu = mu;
            I.ch = phx_ch_params{dw[0], dw[1], dw[2], -b, t1[0].coeff, dt};
            plan_name_ = "fused CH Euler multi-step (phx_ch_euler_steps)";
        }
    }
}

Solver::~Solver() = default;

void Solver::setBulkPrecision(Precision p) {
    if (p == Precision::FP32 && impl_->plan == Impl::GENERIC)
        throw std::invalid_argument("Solver::setBulkPrecision: FP32 bulk needs a fused plan; this equation runs the "
                                    "generic FP64 path");
    precision_ = p;
}

// ------------------------------------------------------------- generic ----
// prev <- curr, enqueued on the null stream (advanceTimeLevelGPU without its
// host wait; the Solver waits once per call).
static void advanceLevel(ScalarField& u) {
    hipOk(hipMemcpyAsync(u.d_prev, u.d_curr, u.storedSize * sizeof(double), hipMemcpyDeviceToDevice, nullptr),
          "Solver: time level");
}

// The list applied object by object, or -- when it is a table of standard
// conditions in axis order (what bcTable accepts) -- as one phx_bc_fill,
// which leaves the same values in one launch.
static void applyBCs(const std::vector<BoundaryCondition*>& bcs, ScalarField& f) {
    phx_grid g = detail::gridOf(f);
    if (!bcs.empty() && detail::bcTable(bcs, f.mesh, g)) {
        detail::throwPhx(phx_bc_fill(&g, PHX_F64, f.d_curr, nullptr), "Solver: BCs");
        return;
    }
    for (auto* bc : bcs) bc->applyOnGPU(f);
}

static void genericEuler(Equation& eq, const std::vector<BoundaryCondition*>& bcs, ScalarField& rhs, double dt) {
    applyBCs(bcs, eq.unknown);
    eq.enqueueRHS(rhs);
    detail::throwPhx(phx_axpy_f64(eq.unknown.d_curr, rhs.d_curr, dt, eq.unknown.storedSize, nullptr), "Euler update");
    advanceLevel(eq.unknown);
}

static void genericRK4(Equation& eq, const std::vector<BoundaryCondition*>& bcs, ScalarField* k[4], ScalarField& tmp,
                       double dt) {
    ScalarField& phi = eq.unknown;
    const std::size_t n = phi.storedSize;
    const double shift[3] = {dt * 0.5, dt * 0.5, dt};
    applyBCs(bcs, phi);
    eq.enqueueRHS(*k[0]);
    for (int s = 1; s < 4; ++s) {
        detail::throwPhx(phx_rk4_tmp_f64(tmp.d_curr, phi.d_curr, k[s - 1]->d_curr, shift[s - 1], n, nullptr), "RK4 tmp");
        applyBCs(bcs, tmp);
        double* saved = phi.d_curr;
        phi.d_curr = tmp.d_curr;
        try {
            eq.enqueueRHS(*k[s]);
        } catch (...) {
            phi.d_curr = saved;
            throw;
        }
        phi.d_curr = saved;
    }
    detail::throwPhx(phx_rk4_update_f64(phi.d_curr, k[0]->d_curr, k[1]->d_curr, k[2]->d_curr, k[3]->d_curr, dt / 6.0, n,
                                        nullptr),
                     "RK4 update");
    advanceLevel(phi);
}

// -------------------------------------------------------------- advance ----
void Solver::advanceN(int n) {
    Impl& I = *impl_;
    if (n <= 0) return;
    if (I.plan == Impl::GENERIC) {
        if (multiStep_) {
            if (I.stepScratch.empty())
                for (auto& s : steps_)
                    I.stepScratch.push_back(s.type == EquationType::TRANSIENT
                                                ? scratchLike(s.equation->unknown, s.equation->name + "_rhs")
                                                : nullptr);
            for (int it = 0; it < n; ++it) {
                for (std::size_t q = 0; q < steps_.size(); ++q) {
                    auto& s = steps_[q];
                    applyBCs(s.bcs, *s.sourceField);
                    if (s.type == EquationType::STEADY) {
                        s.equation->enqueueRHS(s.equation->unknown);
                    } else {
                        s.equation->enqueueRHS(*I.stepScratch[q]);
                        detail::throwPhx(phx_axpy_f64(s.equation->unknown.d_curr, I.stepScratch[q]->d_curr, dt,
                                                      s.equation->unknown.storedSize, nullptr),
                                         "multi-step update");
                    }
                }
                for (auto& s : steps_)
                    if (s.type == EquationType::TRANSIENT) advanceLevel(s.equation->unknown);
            }
            hipOk(hipDeviceSynchronize(), "Solver::advance");
            return;
        }
        if (!I.rhs) {
            I.rhs = scratchLike(equation_.unknown, equation_.unknown.name + "_rhs");
            if (use_rk4_) {
                I.k1 = scratchLike(equation_.unknown, "_k1");
                I.k2 = scratchLike(equation_.unknown, "_k2");
                I.k3 = scratchLike(equation_.unknown, "_k3");
                I.k4 = scratchLike(equation_.unknown, "_k4");
                I.tmp = scratchLike(equation_.unknown, "_phi_tmp");
            }
        }
        for (int it = 0; it < n; ++it) {
            if (use_rk4_) {
                ScalarField* k[4] = {I.k1.get(), I.k2.get(), I.k3.get(), I.k4.get()};
                genericRK4(equation_, bcs_, k, *I.tmp, dt);
            } else {
                genericEuler(equation_, bcs_, *I.rhs, dt);
            }
        }
        // the steps run in order on the null stream; one host wait per call
        // (the reference waits after every update)
        hipOk(hipDeviceSynchronize(), "Solver::advance");
        return;
    }

    // fused plans
    I.ch.dt = dt;
    const bool f32 = precision_ == Precision::FP32;
    const int dtype = f32 ? PHX_F32 : PHX_F64;
    ScalarField& c = *I.c;
    const std::size_t N = c.storedSize;
    if (f32 && !I.c32) {
        hipOk(hipMalloc((void**)&I.c32, N * sizeof(float)), "Solver FP32 state");
        if (I.plan == Impl::CH_EULER) hipOk(hipMalloc((void**)&I.mu32, N * sizeof(float)), "Solver FP32 mu");
    }
    void* cbuf = c.d_curr;
    void* mubuf = I.mu ? I.mu->d_curr : nullptr;
    if (f32) {
        detail::throwPhx(phx_convert_f64_to_f32(I.c32, c.d_curr, N, nullptr), "Solver FP32 in");
        cbuf = I.c32;
        if (I.plan == Impl::CH_EULER) {
            detail::throwPhx(phx_convert_f64_to_f32(I.mu32, I.mu->d_curr, N, nullptr), "Solver FP32 in");
            mubuf = I.mu32;
        }
    }
    if (I.plan == Impl::CH_RK4) {
        if (!I.work) {
            std::size_t bytes = std::max(phx_ch_rk4_workspace_bytes(&I.grid, PHX_F64), (std::size_t)1);
            hipOk(hipMalloc(&I.work, bytes), "Solver RK4 workspace");
        }
        detail::throwPhx(phx_ch_rk4_steps(&I.grid, dtype, cbuf, I.work, &I.ch, n, nullptr), "Solver::advance");
    } else {
        detail::throwPhx(phx_ch_euler_steps(&I.grid, dtype, cbuf, mubuf, &I.ch, n, nullptr), "Solver::advance");
    }
    if (f32) {
        detail::throwPhx(phx_convert_f32_to_f64(c.d_curr, I.c32, N, nullptr), "Solver FP32 out");
        if (I.plan == Impl::CH_EULER)
            detail::throwPhx(phx_convert_f32_to_f64(I.mu->d_curr, I.mu32, N, nullptr), "Solver FP32 out");
    }
    hipOk(hipDeviceSynchronize(), "Solver::advance");
    c.advanceTimeLevelGPU();
}

void Solver::advance() {
    advanceN(1);
    ++step;
    time += dt;
}

void Solver::run(int nSteps, int callbackEvery, std::function<void(const Solver&)> callback) {
    int done = 0;
    while (done < nSteps) {
        int chunk = nSteps - done;
        if (callback && callbackEvery > 0) {
            int toNext = callbackEvery - (step % callbackEvery);
            chunk = std::min(chunk, toNext);
        }
        advanceN(chunk);
        for (int q = 0; q < chunk; ++q) {
            ++step;
            time += dt;
        }
        done += chunk;
        if (callback && callbackEvery > 0 && step % callbackEvery == 0) callback(*this);
    }
}

// ------------------------------------------------------------------ CPU ----
// Host twin of the generic path. RK4 applies the BCs to every stage state (the
// device semantics); the reference's host RK4 refreshes the ghosts of phi
// instead (Solver.cu:281-283), which leaves stages 2-4 reading stale ghosts.
void Solver::advanceCPU() {
    if (multiStep_) {
        Impl& I = *impl_;
        if (I.stepScratch.empty())
            for (auto& s : steps_)
                I.stepScratch.push_back(s.type == EquationType::TRANSIENT
                                            ? std::make_unique<ScalarField>(s.equation->unknown.mesh, "_rhs",
                                                                            s.equation->unknown.ghost)
                                            : nullptr);
        for (std::size_t q = 0; q < steps_.size(); ++q) {
            auto& s = steps_[q];
            for (auto* bc : s.bcs) bc->applyOnCPU(*s.sourceField);
            if (s.type == EquationType::STEADY) {
                s.equation->computeRHSCPU(s.equation->unknown);
            } else {
                s.equation->computeRHSCPU(*I.stepScratch[q]);
                auto& u = s.equation->unknown.curr;
                const auto& r = I.stepScratch[q]->curr;
                for (std::size_t i = 0; i < u.size(); ++i) u[i] += dt * r[i];
            }
        }
        for (auto& s : steps_)
            if (s.type == EquationType::TRANSIENT) s.equation->unknown.advanceTimeLevelCPU();
    } else {
        ScalarField& phi = equation_.unknown;
        const std::size_t N = phi.storedSize;
        ScalarField rhs(phi.mesh, "_rhs_cpu", phi.ghost);
        if (!use_rk4_) {
            for (auto* bc : bcs_) bc->applyOnCPU(phi);
            equation_.computeRHSCPU(rhs);
            for (std::size_t i = 0; i < N; ++i) phi.curr[i] += dt * rhs.curr[i];
        } else {
            std::vector<double> k[4];
            ScalarField tmp(phi.mesh, "_tmp_cpu", phi.ghost);
            const double shift[3] = {dt * 0.5, dt * 0.5, dt};
            for (auto* bc : bcs_) bc->applyOnCPU(phi);
            equation_.computeRHSCPU(rhs);
            k[0] = rhs.curr;
            for (int s = 1; s < 4; ++s) {
                for (std::size_t i = 0; i < N; ++i) tmp.curr[i] = phi.curr[i] + shift[s - 1] * k[s - 1][i];
                for (auto* bc : bcs_) bc->applyOnCPU(tmp);
                std::swap(phi.curr, tmp.curr);
                equation_.computeRHSCPU(rhs);
                std::swap(phi.curr, tmp.curr);
                k[s] = rhs.curr;
            }
            const double dt6 = dt / 6.0;
            for (std::size_t i = 0; i < N; ++i)
                phi.curr[i] += dt6 * (k[0][i] + 2.0 * k[1][i] + 2.0 * k[2][i] + k[3][i]);
        }
        phi.advanceTimeLevelCPU();
    }
    ++step;
    time += dt;
}

}  // namespace PhiX
