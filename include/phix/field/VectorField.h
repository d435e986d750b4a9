#pragma once
// SoA vector field of N ScalarFields; drop-in for the reference VectorField
// (include/field/VectorField.h:32-143). Components are name_x/_y/_z for N == 3,
// name_0.. otherwise.
#include <string>
#include <vector>

#include "field/ScalarField.h"
#include "mesh/Mesh.h"

namespace PhiX {

class VectorField {
public:
    explicit VectorField(const Mesh& mesh, const std::string& name, int nComponents, int ghost = 1);
    VectorField(const VectorField&) = delete;
    VectorField& operator=(const VectorField&) = delete;
    VectorField(VectorField&& other) noexcept;
    VectorField& operator=(VectorField&& other) noexcept;
    ~VectorField() = default;

    ScalarField& operator[](int c);
    const ScalarField& operator[](int c) const;
    int nComponents() const { return static_cast<int>(components_.size()); }

    ScalarField& x() { return (*this)[0]; }
    ScalarField& y() { return (*this)[1]; }
    ScalarField& z() { return (*this)[2]; }
    const ScalarField& x() const { return (*this)[0]; }
    const ScalarField& y() const { return (*this)[1]; }
    const ScalarField& z() const { return (*this)[2]; }

    const Mesh& mesh;
    int ghost;
    std::string name;

    void fill(double value);
    void fillCurr(double value);
    void fillPrev(double value);
    void advanceTimeLevelCPU();
    void advanceTimeLevelGPU();

    bool deviceAllocated() const;
    void allocDevice();
    void freeDevice();
    void uploadCurrToDevice() const;
    void uploadPrevToDevice() const;
    void uploadAllToDevice() const;
    void downloadCurrFromDevice();
    void downloadPrevFromDevice();
    void downloadAllFromDevice();

    void write(const std::string& path, FieldFormat fmt = FieldFormat::BINARY) const;
    static VectorField readFromFile(const Mesh& mesh, const std::string& path, int ghost = 1);
    void print() const;

private:
    std::vector<ScalarField> components_;
    static std::string componentName(const std::string& baseName, int c, int nComp);
};

}  // namespace PhiX
