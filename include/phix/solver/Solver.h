#pragma once
// Explicit time stepping; drop-in for include/solver/Solver.h:17-162.
//
// Construction plans the step once. If the equation is one of the forms the
// fused gfx950 kernels implement, advance()/run() go through the C ABI
// (include/phix_amd.h):
//   * single equation, RK4, RHS = M*lap(pw(c, f') - kappa*lap(c), bcs) with the
//     double-well f' and bcs equal to the solver's BCs      -> phx_ch_rk4_steps
//   * multi-step {c, bcs, mu = pw(c, f') - kappa*lap(c), STEADY},
//                {mu, bcs, dc/dt = M*lap(mu), TRANSIENT}    -> phx_ch_euler_steps
// Every other equation runs the generic path: one kernel per Term, the
// reference's RK4 pointer swap and BC launches, with identical results.
// plan() names the path chosen.
//
// Not in the reference: setBulkPrecision(FP32) runs a fused plan on an FP32
// copy of the unknown (refreshed from d_curr at the start of each
// advance()/run() and written back to d_curr/d_prev in double at the end).
#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "boundary/BoundaryCondition.h"
#include "equation/Equation.h"
#include "field/ScalarField.h"

namespace PhiX {

enum class TimeScheme { EULER, RK4 };
enum class EquationType { TRANSIENT, STEADY };
enum class Precision { FP64, FP32 };

struct SolverStep {
    ScalarField* sourceField;
    std::vector<BoundaryCondition*> bcs;
    Equation* equation;
    EquationType type = EquationType::TRANSIENT;
};

class Solver {
public:
    Solver(Equation& equation, std::vector<BoundaryCondition*> bcs, double dt,
           TimeScheme scheme = TimeScheme::EULER);
    Solver(std::vector<SolverStep> steps, double dt, TimeScheme scheme = TimeScheme::EULER);
    ~Solver();

    Solver(const Solver&) = delete;
    Solver& operator=(const Solver&) = delete;

    double dt;
    TimeScheme scheme;
    int step = 0;
    double time = 0.0;

    void advance();
    void advanceCPU();
    void run(int nSteps, int callbackEvery = 0, std::function<void(const Solver&)> callback = nullptr);

    const ScalarField& unknown() const { return equation_.unknown; }
    ScalarField& unknown() { return equation_.unknown; }
    const Equation& equation() const { return equation_; }
    const std::vector<BoundaryCondition*>& bcs() const { return bcs_; }

    // Extensions (no reference counterpart).
    void setBulkPrecision(Precision p);
    Precision bulkPrecision() const { return precision_; }
    const std::string& plan() const { return plan_name_; }

private:
    struct Impl;
    Equation& equation_;
    std::vector<BoundaryCondition*> bcs_;
    bool use_rk4_ = false;
    bool multiStep_ = false;
    std::vector<SolverStep> steps_;
    Precision precision_ = Precision::FP64;
    std::string plan_name_;
    std::unique_ptr<Impl> impl_;

    void advanceN(int n);   // n steps on the device, unknown complete on return
};

}  // namespace PhiX
