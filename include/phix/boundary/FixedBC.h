#pragma once
// "Dirichlet" as the reference implements it: every ghost layer on the chosen
// side(s) is set to `value` (a constant fill, include/boundary/FixedBC.h:11-16,
// src/boundary/Boundary.cu:166-195).
#include "boundary/BoundaryCondition.h"

namespace PhiX {

class FixedBC : public BoundaryCondition {
public:
    double value;

    FixedBC(Axis axis, Side side, double value);

    using BoundaryCondition::applyOnCPU;
    using BoundaryCondition::applyOnGPU;

    void applyOnCPU(ScalarField& f) const override;
    void applyOnGPU(ScalarField& f) const override;
    int kind() const override { return 3; }
    double fixedValue() const override { return value; }
};

}  // namespace PhiX
