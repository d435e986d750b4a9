#pragma once
// Periodic condition -- ghost layers wrap from the opposite physical end; the
// side is always BOTH. Drop-in for include/boundary/PeriodicBC.h.
#include "boundary/BoundaryCondition.h"

namespace PhiX {

class PeriodicBC : public BoundaryCondition {
public:
    explicit PeriodicBC(Axis axis);

    using BoundaryCondition::applyOnCPU;
    using BoundaryCondition::applyOnGPU;

    void applyOnCPU(ScalarField& f) const override;
    void applyOnGPU(ScalarField& f) const override;
    int kind() const override { return 1; }
};

}  // namespace PhiX
