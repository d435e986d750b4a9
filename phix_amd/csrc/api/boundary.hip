// boundary.hip -- PeriodicBC / NoFluxBC / FixedBC of the drop-in C++ API.
// applyOnGPU enqueues phx_bc_apply on the null stream (one launch per BC
// object, as in the reference, src/boundary/Boundary.cu:224-351); applyOnCPU
// runs the same face rule on the host arrays.
#include <stdexcept>
#include <string>

#include "../../../include/phix_amd.h"
#include "boundary/FixedBC.h"
#include "boundary/NoFluxBC.h"
#include "boundary/PeriodicBC.h"

namespace PhiX {
namespace detail {

phx_grid gridOf(const ScalarField& f) {
    phx_grid g{};
    g.dim = f.mesh.dim;
    for (int a = 0; a < 3; ++a) {
        g.n[a] = f.mesh.n[a];
        g.d[a] = f.mesh.d[a];
    }
    g.ghost = f.ghost;
    return g;
}

void throwPhx(int rc, const std::string& who) {
    if (rc == PHX_OK) return;
    std::string msg = who + ": " + phx_last_error();
    if (rc == PHX_ERR_INVALID) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

// Host twin of the face kernels: every (face0, face1) position over the full
// stored extent of the other two axes, ghost layers l = 1..g.
static void applyFaceCPU(ScalarField& f, int axis, int kind, bool lo, bool hi, double value) {
    if (f.curr.size() != f.storedSize) throw std::runtime_error("BC applyOnCPU: field has no host storage");
    const long long sx = f.storedDims[0], sxy = sx * f.storedDims[1];
    const long long st = axis == 0 ? 1 : (axis == 1 ? sx : sxy);
    const int other[2] = {axis == 0 ? 1 : 0, axis == 2 ? 1 : 2};
    const long long ostride[2] = {other[0] == 0 ? 1 : (other[0] == 1 ? sx : sxy),
                                  other[1] == 0 ? 1 : (other[1] == 1 ? sx : sxy)};
    const int g = f.ghost, n = f.mesh.n[axis];
    double* d = f.curr.data();
    for (int b = 0; b < f.storedDims[other[1]]; ++b)
        for (int a = 0; a < f.storedDims[other[0]]; ++a) {
            const long long base = a * ostride[0] + b * ostride[1];
            for (int l = 1; l <= g; ++l) {
                const long long glo = base + (g - l) * st, ghi = base + (g + n + l - 1) * st;
                if (kind == PHX_BC_PERIODIC) {
                    d[glo] = d[base + (g + n - l) * st];
                    d[ghi] = d[base + (g + l - 1) * st];
                } else if (kind == PHX_BC_NOFLUX) {
                    if (lo) d[glo] = d[base + g * st];
                    if (hi) d[ghi] = d[base + (g + n - 1) * st];
                } else {
                    if (lo) d[glo] = value;
                    if (hi) d[ghi] = value;
                }
            }
        }
}

static void applyFaceGPU(ScalarField& f, int axis, int kind, Side side, double value, const char* who) {
    if (!f.deviceAllocated()) throw std::runtime_error(std::string(who) + "::applyOnGPU: device not allocated");
    phx_grid g = gridOf(f);
    int s = side == Side::LOW ? PHX_SIDE_LOW : (side == Side::HIGH ? PHX_SIDE_HIGH : PHX_SIDE_BOTH);
    throwPhx(phx_bc_apply(&g, PHX_F64, f.d_curr, axis, kind, s, value, nullptr), who);
}

}  // namespace detail

PeriodicBC::PeriodicBC(Axis a) : BoundaryCondition(a, Side::BOTH) {}
void PeriodicBC::applyOnCPU(ScalarField& f) const {
    detail::applyFaceCPU(f, (int)axis, PHX_BC_PERIODIC, true, true, 0.0);
}
void PeriodicBC::applyOnGPU(ScalarField& f) const {
    detail::applyFaceGPU(f, (int)axis, PHX_BC_PERIODIC, Side::BOTH, 0.0, "PeriodicBC");
}

NoFluxBC::NoFluxBC(Axis a, Side s) : BoundaryCondition(a, s) {}
void NoFluxBC::applyOnCPU(ScalarField& f) const {
    detail::applyFaceCPU(f, (int)axis, PHX_BC_NOFLUX, side != Side::HIGH, side != Side::LOW, 0.0);
}
void NoFluxBC::applyOnGPU(ScalarField& f) const {
    detail::applyFaceGPU(f, (int)axis, PHX_BC_NOFLUX, side, 0.0, "NoFluxBC");
}

FixedBC::FixedBC(Axis a, Side s, double v) : BoundaryCondition(a, s), value(v) {}
void FixedBC::applyOnCPU(ScalarField& f) const {
    detail::applyFaceCPU(f, (int)axis, PHX_BC_FIXED, side != Side::HIGH, side != Side::LOW, value);
}
void FixedBC::applyOnGPU(ScalarField& f) const {
    detail::applyFaceGPU(f, (int)axis, PHX_BC_FIXED, side, value, "FixedBC");
}

}  // namespace PhiX
