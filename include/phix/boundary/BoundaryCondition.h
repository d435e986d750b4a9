#pragma once
// Boundary-condition interface; drop-in for include/boundary/BoundaryCondition.h:12-43.
#include "field/ScalarField.h"
#include "field/VectorField.h"

namespace PhiX {

enum class Axis { X = 0, Y = 1, Z = 2 };
enum class Side { LOW, HIGH, BOTH };

class BoundaryCondition {
public:
    Axis axis;
    Side side;

    BoundaryCondition(Axis axis, Side side) : axis(axis), side(side) {}
    virtual ~BoundaryCondition() = default;

    virtual void applyOnCPU(ScalarField& f) const = 0;
    virtual void applyOnGPU(ScalarField& f) const = 0;

    virtual void applyOnCPU(VectorField& vf) const {
        for (int c = 0; c < vf.nComponents(); ++c) applyOnCPU(vf[c]);
    }
    virtual void applyOnGPU(VectorField& vf) const {
        for (int c = 0; c < vf.nComponents(); ++c) applyOnGPU(vf[c]);
    }

    // Kind as the fused kernels see it (phx_bc_kind of include/phix_amd.h) and
    // the FIXED constant. Not part of the reference interface; the Solver
    // planner uses it to turn a BC list into per-face index maps.
    virtual int kind() const = 0;
    virtual double fixedValue() const { return 0.0; }
};

}  // namespace PhiX
