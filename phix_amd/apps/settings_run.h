#pragma once
// Pieces shared by the config-driven solver programs in this folder: the mesh
// block of a settings file and the time loop with its print/write cadence.
#include <algorithm>
#include <iostream>
#include <vector>

#include "IO/Json.h"
#include "IO/OutputWriter.h"
#include "field/ScalarField.h"
#include "mesh/Mesh.h"
#include "solver/Solver.h"

namespace PhiX {
namespace apps {

// "mesh": {nx, ny, dx, dy, x0, y0} as the reference's 2D drivers read it
// (applications/solvers/Cahn-Hillard_double-well/2D/Cahn-Hillard_double-well.cu:40-49).
// Optional nz (with dz, z0) makes the mesh 3D; the reference's drivers are 2D.
inline Mesh meshFrom(const IO::Json& m) {
    const int nx = m["nx"], ny = m["ny"];
    const double dx = m["dx"], dy = m["dy"], x0 = m["x0"], y0 = m["y0"];
    if (m.contains("nz") && m["nz"].get<int>() > 1)
        return Mesh::makeUniform3D(CoordSys::CARTESIAN, nx, dx, x0, ny, dy, y0, m["nz"], m.value("dz", dx),
                                   m.value("z0", 0.0));
    return Mesh::makeUniform2D(CoordSys::CARTESIAN, nx, dx, x0, ny, dy, y0);
}

// "initialize": {"bulk_precision": "FP32"} (optional, absent in the
// reference's settings) runs the fused kernels on FP32 copies of the fields.
inline void applyPrecision(Solver& s, const IO::Json& init) {
    const std::string p = init.value("bulk_precision", "FP64");
    if (p == "FP32") s.setBulkPrecision(Precision::FP32);
    else if (p != "FP64") throw std::runtime_error("bulk_precision must be \"FP64\" or \"FP32\", got \"" + p + "\"");
}

// Advances from solver.step to nSteps (the last step number, as the reference's
// loop bound: Cahn-Hillard_double-well.cu:127-135). After every step that is a
// multiple of the print interval the progress line is printed, after every
// multiple of the write interval the fields are written. The steps in between
// go to the device as one Solver::run, so the host waits only at those events;
// the stepping and its results are the same as one advance() per step.
inline void timeLoop(Solver& solver, IO::OutputWriter& writer, const std::vector<ScalarField*>& out, int nSteps) {
    auto nextMultiple = [&](int every) { return (solver.step / every + 1) * every; };
    while (solver.step < nSteps) {
        const int stop = std::min({nextMultiple(writer.printInterval), nextMultiple(writer.writeInterval), nSteps});
        solver.run(stop - solver.step);
        if (writer.shouldPrint(solver.step)) writer.printProgress(solver.step, solver.time);
        if (writer.shouldWrite(solver.step))
            for (ScalarField* f : out) writer.writeFields(*f, solver.step, solver.time);
    }
}

}  // namespace apps
}  // namespace PhiX
