#pragma once
// Ghost-padded FP64 scalar field with host and device copies; drop-in for the
// reference ScalarField (include/field/ScalarField.h:30-169).
//   storedDims[a] = n[a] + 2*ghost,  index = (i+g) + sx*((j+g) + sy*(k+g))
// Host curr/prev always exist; device d_curr/d_prev appear with allocDevice().
#include <cstddef>
#include <string>
#include <vector>

#include "IO/FieldFormat.h"
#include "mesh/Mesh.h"

namespace PhiX {

class ScalarField {
public:
    std::string name;
    const Mesh& mesh;
    int ghost;
    int storedDims[3];
    std::size_t storedSize;

    std::vector<double> curr;
    std::vector<double> prev;

    double* d_curr = nullptr;
    double* d_prev = nullptr;

    explicit ScalarField(const Mesh& mesh, const std::string& name, int ghost = 1);

    // Non-owning view of an external device buffer (d_curr only); the
    // destructor does not free it. Used to run BCs on scratch buffers.
    static ScalarField makeShell(const Mesh& mesh, int ghost, double* d_buf, const std::string& name = "shell");

    ScalarField(const ScalarField&) = delete;
    ScalarField& operator=(const ScalarField&) = delete;
    ScalarField(ScalarField&& other) noexcept;
    ScalarField& operator=(ScalarField&& other) noexcept;
    ~ScalarField();

    inline int index(int i, int j, int k) const {
        return (i + ghost) + storedDims[0] * ((j + ghost) + storedDims[1] * (k + ghost));
    }
    inline int index(int i, int j) const { return index(i, j, 0); }
    inline int index(int i) const { return index(i, 0, 0); }

    void fill(double value);
    void fillCurr(double value);
    void fillPrev(double value);

    void advanceTimeLevelCPU();
    void advanceTimeLevelGPU();

    bool deviceAllocated() const { return d_curr != nullptr; }
    void allocDevice();
    void freeDevice();
    void uploadCurrToDevice() const;
    void uploadPrevToDevice() const;
    void uploadAllToDevice() const;
    void downloadCurrFromDevice();
    void downloadPrevFromDevice();
    void downloadAllFromDevice();

    bool ownsDeviceMemory() const { return ownsDeviceMemory_; }

    void write(const std::string& path, FieldFormat fmt = FieldFormat::BINARY) const;
    static ScalarField readFromFile(const Mesh& mesh, const std::string& path, int ghost = 1);
    void print() const;

private:
    bool ownsDeviceMemory_ = true;
};

using Field = ScalarField;

}  // namespace PhiX
