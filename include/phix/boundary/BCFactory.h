#pragma once
// Boundary conditions from a settings block; drop-in for
// include/boundary/BCFactory.h:
//   { "x_min": "Periodic" | "NoFlux", "x_max": ..., "y_min": ..., "y_max": ...,
//     "z_min": ..., "z_max": ... }   (y and z pairs optional)
// Periodic must be set on both sides of an axis (else std::runtime_error);
// other types are rejected with std::runtime_error, as in BCFactory.cpp:13-74.
#include "IO/Json.h"
#include "boundary/BoundaryCondition.h"

#include <memory>
#include <vector>

namespace PhiX {

struct BCSet {
    std::vector<std::unique_ptr<BoundaryCondition>> storage;   // owns the conditions
    std::vector<BoundaryCondition*> ptrs;                      // the list a Solver takes
};

BCSet buildBCs(const IO::Json& bc_config);

}  // namespace PhiX
