#pragma once
// d(unknown)/dt = RHS; drop-in for include/equation/Equation.h:27-58.
// computeRHS zeroes the whole stored rhs and accumulates the terms in order on
// the device; it returns with the result complete (device synchronised).
#include <string>
#include <vector>

#include "equation/Term.h"
#include "field/ScalarField.h"

namespace PhiX {

class Equation {
public:
    std::string name;
    ScalarField& unknown;
    std::vector<ScalarField*> auxFields;

    explicit Equation(ScalarField& unknown, const std::string& name = "");

    void setRHS(const RHSExpr& expr);
    void setRHS(const Term& t);

    void computeRHS(ScalarField& rhs) const;
    void computeRHSCPU(ScalarField& rhs) const;

    bool hasRHS() const { return !rhs_expr_.terms.empty(); }

    // The expression as set (read by the Solver planner; not in the reference).
    const RHSExpr& rhsExpr() const { return rhs_expr_; }

    // computeRHS without the final host wait: the work is enqueued in order on
    // the null stream (used by the Solver's steps, which wait once per call;
    // not in the reference).
    void enqueueRHS(ScalarField& rhs) const;

private:
    RHSExpr rhs_expr_;
    mutable ScratchPool scratch_pool_;
};

}  // namespace PhiX
