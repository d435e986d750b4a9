#pragma once
// Explicit time stepping of a VectorEquation; drop-in for
// include/solver/VectorSolver.h:32-94. All components' stage states are swapped
// together so coupled right-hand sides see one consistent stage.
//
// A VectorEquation whose component i is
//   (L*kappa)*lap(eta_i) + pw(eta_i, e - e^3, L) + sum_{j!=i} pw(eta_i, eta_j, a*b*b, -2*L*gamma)
// (terms in this order) with RK4 runs as one fused kernel per stage
// (phx_ac_rk4_steps); anything else runs the generic per-term path.
#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "boundary/BoundaryCondition.h"
#include "equation/VectorEquation.h"
#include "field/VectorField.h"
#include "solver/Solver.h"

namespace PhiX {

class VectorSolver {
public:
    VectorSolver(VectorEquation& equation, std::vector<BoundaryCondition*> bcs, double dt,
                 TimeScheme scheme = TimeScheme::EULER);
    ~VectorSolver();
    VectorSolver(const VectorSolver&) = delete;
    VectorSolver& operator=(const VectorSolver&) = delete;

    double dt;
    TimeScheme scheme;
    int step = 0;
    double time = 0.0;

    void advance();
    void advanceCPU();
    void run(int nSteps, int callbackEvery = 0, std::function<void(const VectorSolver&)> callback = nullptr);

    const VectorField& unknown() const { return equation_.unknown; }
    VectorField& unknown() { return equation_.unknown; }
    const VectorEquation& equation() const { return equation_; }
    const std::vector<BoundaryCondition*>& bcs() const { return bcs_; }

    void setBulkPrecision(Precision p);
    const std::string& plan() const { return plan_name_; }

private:
    struct Impl;
    VectorEquation& equation_;
    std::vector<BoundaryCondition*> bcs_;
    bool use_rk4_;
    Precision precision_ = Precision::FP64;
    std::string plan_name_;
    std::unique_ptr<Impl> impl_;
    void advanceN(int n);
};

}  // namespace PhiX
