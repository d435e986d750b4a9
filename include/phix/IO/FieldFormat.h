#pragma once
// On-disk formats of ScalarField::write (reference: include/IO/FieldFormat.h).
namespace PhiX {
enum class FieldFormat { BINARY, DAT, VTS };
}  // namespace PhiX
