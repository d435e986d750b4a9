#pragma once
// Field file IO; drop-in for include/IO/FieldIO.h (free functions of
// PhiX::IO). Formats: BINARY (.field: text header, then the physical cells as
// raw FP64, x fastest), DAT (text, "x y z value" per cell), VTS (VTK XML
// StructuredGrid, cell data; write only).
#include "IO/FieldFormat.h"
#include "mesh/Mesh.h"

#include <string>

namespace PhiX {
class ScalarField;
class VectorField;
namespace IO {

void writeField(const ScalarField& f, const std::string& path, FieldFormat fmt = FieldFormat::BINARY);
ScalarField readScalarField(const Mesh& mesh, const std::string& path, int ghost = 1);
// Reads a BINARY or DAT file into an existing field's physical cells (curr);
// the file's dimensions must match the field's mesh.
void readField(ScalarField& f, const std::string& path);

// "initial_field" -> 0; otherwise a step number, resolved to the step of the
// output/<refFieldName>_<step>.field file nearest to it (FieldIO.cpp:750-782).
int resolveStartStep(const std::string& startFrom, const std::string& refFieldName = "c");
// Step 0: settings/initial_field/<name>.field; else output/<name>_<step>.field.
void initField(ScalarField& f, int startStep);

void writeField(const VectorField& vf, const std::string& path, FieldFormat fmt = FieldFormat::BINARY);
VectorField readVectorField(const Mesh& mesh, const std::string& path, int ghost = 1);
void readField(VectorField& vf, const std::string& path);

}  // namespace IO
}  // namespace PhiX
