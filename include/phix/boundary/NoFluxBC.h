#pragma once
// Zero-gradient (Neumann) condition -- every ghost layer on the chosen side(s)
// copies the nearest physical cell. Drop-in for include/boundary/NoFluxBC.h.
#include "boundary/BoundaryCondition.h"

namespace PhiX {

class NoFluxBC : public BoundaryCondition {
public:
    NoFluxBC(Axis axis, Side side = Side::BOTH);

    using BoundaryCondition::applyOnCPU;
    using BoundaryCondition::applyOnGPU;

    void applyOnCPU(ScalarField& f) const override;
    void applyOnGPU(ScalarField& f) const override;
    int kind() const override { return 2; }
};

}  // namespace PhiX
