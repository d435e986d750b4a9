#pragma once
// Uniform structured grid, drop-in for the reference Mesh
// (include/mesh/Mesh.h:14-95, src/mesh/Mesh.cpp:39-111): dim 1..3, unused axes
// have n = 1, cell-centre coordinates, x-fastest physical index.
#include <cstddef>
#include <stdexcept>
#include <string>

namespace PhiX {

enum class CoordSys { CARTESIAN, CYLINDRICAL, SPHERICAL };

class Mesh {
public:
    int dim;
    CoordSys coordSys;
    int n[3];
    double d[3];
    double origin[3];

    Mesh();
    Mesh(int dim, CoordSys cs, int nx, double dx, double x0, int ny, double dy, double y0, int nz, double dz,
         double z0);

    static Mesh makeUniform1D(CoordSys cs, int nx, double dx, double x0 = 0.0);
    static Mesh makeUniform2D(CoordSys cs, int nx, double dx, double x0, int ny, double dy, double y0);
    static Mesh makeUniform3D(CoordSys cs, int nx, double dx, double x0, int ny, double dy, double y0, int nz,
                              double dz, double z0);

    static Mesh readFromFile(const std::string& path);
    void write(const std::string& path) const;
    void print() const;

    inline std::size_t totalSize() const { return static_cast<std::size_t>(n[0]) * n[1] * n[2]; }
    inline double coord(int axis, int i) const { return origin[axis] + (i + 0.5) * d[axis]; }
    inline int index(int i, int j, int k) const { return i + n[0] * (j + n[1] * k); }
    inline int index(int i) const { return i; }
    inline int index(int i, int j) const { return i + n[0] * j; }

    void validate() const;
    bool isValid() const noexcept;

private:
    void init(int dim_, CoordSys cs_, int nx, double dx, double x0, int ny, double dy, double y0, int nz,
              double dz, double z0);
};

std::string coordSysToString(CoordSys cs);
CoordSys stringToCoordSys(const std::string& s);

}  // namespace PhiX
