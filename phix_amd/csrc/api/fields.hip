// fields.hip -- Mesh, ScalarField, VectorField of the drop-in C++ API.
//
// Device memory: hipMalloc'd FP64 stored arrays, zero-filled on allocation
// (the reference's allocDevice contract, src/field/ScalarField.cu:162-170).
// advanceTimeLevelGPU copies curr -> prev on the device.
//
// On-disk BINARY format (.field / .vfield) is the reference's: a text header
// ("# PhiX ScalarField", name, nx ny nz, ghost, "---") followed by the
// physical cells as raw FP64, x fastest (include/field/ScalarField.h:118-140),
// so files interoperate with the reference's Python readers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cctype>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <limits>
#include <sstream>
#include <stdexcept>
#include <string>

#include "field/ScalarField.h"
#include "field/VectorField.h"
#include "mesh/Mesh.h"

namespace PhiX {

#define PHIX_HIP_CHECK(call)                                                                          \
    do {                                                                                              \
        hipError_t e_ = (call);                                                                       \
        if (e_ != hipSuccess)                                                                         \
            throw std::runtime_error(std::string("HIP error in ") + __FILE__ + " line " +             \
                                     std::to_string(__LINE__) + ": " + hipGetErrorString(e_));        \
    } while (0)

// ------------------------------------------------------------------ Mesh ----
std::string coordSysToString(CoordSys cs) {
    switch (cs) {
        case CoordSys::CYLINDRICAL: return "CYLINDRICAL";
        case CoordSys::SPHERICAL: return "SPHERICAL";
        default: return "CARTESIAN";
    }
}

CoordSys stringToCoordSys(const std::string& s) {
    std::string u(s);
    for (auto& ch : u) ch = (char)std::toupper((unsigned char)ch);
    if (u == "CARTESIAN") return CoordSys::CARTESIAN;
    if (u == "CYLINDRICAL") return CoordSys::CYLINDRICAL;
    if (u == "SPHERICAL") return CoordSys::SPHERICAL;
    throw std::invalid_argument("Unknown CoordSys: " + s);
}

void Mesh::init(int dim_, CoordSys cs_, int nx, double dx, double x0, int ny, double dy, double y0, int nz,
                double dz, double z0) {
    dim = dim_;
    coordSys = cs_;
    const int ns[3] = {nx, ny, nz};
    const double ds[3] = {dx, dy, dz}, os[3] = {x0, y0, z0};
    for (int a = 0; a < 3; ++a) {
        n[a] = ns[a];
        d[a] = ds[a];
        origin[a] = os[a];
    }
}

Mesh::Mesh() { init(1, CoordSys::CARTESIAN, 0, 0.0, 0.0, 1, 1.0, 0.0, 1, 1.0, 0.0); }

Mesh::Mesh(int dim_, CoordSys cs, int nx, double dx, double x0, int ny, double dy, double y0, int nz, double dz,
           double z0) {
    init(dim_, cs, nx, dx, x0, ny, dy, y0, nz, dz, z0);
    validate();
}

Mesh Mesh::makeUniform1D(CoordSys cs, int nx, double dx, double x0) {
    return Mesh(1, cs, nx, dx, x0, 1, 1.0, 0.0, 1, 1.0, 0.0);
}
Mesh Mesh::makeUniform2D(CoordSys cs, int nx, double dx, double x0, int ny, double dy, double y0) {
    return Mesh(2, cs, nx, dx, x0, ny, dy, y0, 1, 1.0, 0.0);
}
Mesh Mesh::makeUniform3D(CoordSys cs, int nx, double dx, double x0, int ny, double dy, double y0, int nz, double dz,
                         double z0) {
    return Mesh(3, cs, nx, dx, x0, ny, dy, y0, nz, dz, z0);
}

void Mesh::validate() const {
    if (dim < 1 || dim > 3) throw std::invalid_argument("Mesh: dim must be 1, 2, or 3");
    for (int a = 0; a < 3; ++a) {
        const std::string tag = "Mesh: n[" + std::to_string(a) + "]";
        if (a < dim) {
            if (n[a] <= 0) throw std::invalid_argument(tag + " must be > 0");
            if (!(d[a] > 0.0)) throw std::invalid_argument("Mesh: d[" + std::to_string(a) + "] must be > 0");
        } else if (n[a] != 1) {
            throw std::invalid_argument(tag + " must be 1 for unused axes");
        }
    }
}

bool Mesh::isValid() const noexcept {
    try {
        validate();
    } catch (...) {
        return false;
    }
    return true;
}

void Mesh::write(const std::string& path) const {
    validate();
    std::ofstream os(path);
    if (!os) throw std::runtime_error("Mesh::write: cannot open file: " + path);
    os << "# PhiX Mesh\n"
       << "dim      " << dim << "\n"
       << "coordSys " << coordSysToString(coordSys) << "\n";
    const char* ax[3] = {"x", "y", "z"};
    os.precision(17);
    for (int a = 0; a < 3; ++a)
        os << "n" << ax[a] << " " << n[a] << "  d" << ax[a] << " " << d[a] << "  " << ax[a] << "0 " << origin[a]
           << "\n";
}

Mesh Mesh::readFromFile(const std::string& path) {
    std::ifstream is(path);
    if (!is) throw std::runtime_error("Mesh::readFromFile: cannot open file: " + path);
    int dim_ = 1;
    CoordSys cs = CoordSys::CARTESIAN;
    int nn[3] = {1, 1, 1};
    double dd[3] = {1.0, 1.0, 1.0}, oo[3] = {0.0, 0.0, 0.0};
    std::string line;
    while (std::getline(is, line)) {
        auto h = line.find('#');
        if (h != std::string::npos) line.erase(h);
        std::istringstream ss(line);
        std::string tok;
        while (ss >> tok) {
            if (tok == "dim") ss >> dim_;
            else if (tok == "coordSys") {
                std::string v;
                ss >> v;
                cs = stringToCoordSys(v);
            } else if (tok.size() == 2) {
                int a = tok[0] == 'x' || tok[1] == 'x' ? 0 : (tok[0] == 'y' || tok[1] == 'y' ? 1 : 2);
                if (tok[0] == 'n') ss >> nn[a];
                else if (tok[0] == 'd') ss >> dd[a];
                else if (tok[1] == '0') ss >> oo[a];
            }
        }
    }
    return Mesh(dim_, cs, nn[0], dd[0], oo[0], nn[1], dd[1], oo[1], nn[2], dd[2], oo[2]);
}

void Mesh::print() const {
    std::cout << "=== Mesh ===\n  dim      : " << dim << "\n  coordSys : " << coordSysToString(coordSys) << "\n";
    const char* ax[3] = {"x", "y", "z"};
    for (int a = 0; a < 3; ++a)
        std::cout << "  " << ax[a] << ": n=" << n[a] << "  d=" << d[a] << "  origin=" << origin[a]
                  << (a >= dim ? "  (inactive)" : "") << "\n";
    std::cout << "  totalSize: " << totalSize() << "\n";
}

// ----------------------------------------------------------- ScalarField ----
ScalarField::ScalarField(const Mesh& m, const std::string& nm, int g) : name(nm), mesh(m), ghost(g) {
    if (g < 0) throw std::invalid_argument("ScalarField: ghost must be >= 0");
    storedSize = 1;
    for (int a = 0; a < 3; ++a) {
        storedDims[a] = mesh.n[a] + 2 * g;
        storedSize *= (std::size_t)storedDims[a];
    }
    curr.assign(storedSize, 0.0);
    prev.assign(storedSize, 0.0);
}

ScalarField ScalarField::makeShell(const Mesh& m, int g, double* d_buf, const std::string& nm) {
    ScalarField f(m, nm, g);
    std::vector<double>().swap(f.curr);
    std::vector<double>().swap(f.prev);
    f.d_curr = d_buf;
    f.d_prev = nullptr;
    f.ownsDeviceMemory_ = false;
    return f;
}

ScalarField::ScalarField(ScalarField&& o) noexcept
    : name(std::move(o.name)),
      mesh(o.mesh),
      ghost(o.ghost),
      storedSize(o.storedSize),
      curr(std::move(o.curr)),
      prev(std::move(o.prev)),
      d_curr(o.d_curr),
      d_prev(o.d_prev),
      ownsDeviceMemory_(o.ownsDeviceMemory_) {
    std::copy(o.storedDims, o.storedDims + 3, storedDims);
    o.d_curr = o.d_prev = nullptr;
    o.ownsDeviceMemory_ = false;
}

ScalarField& ScalarField::operator=(ScalarField&& o) noexcept {
    if (this == &o) return *this;
    freeDevice();
    name = std::move(o.name);
    ghost = o.ghost;  // the mesh reference cannot be rebound
    std::copy(o.storedDims, o.storedDims + 3, storedDims);
    storedSize = o.storedSize;
    curr = std::move(o.curr);
    prev = std::move(o.prev);
    d_curr = o.d_curr;
    d_prev = o.d_prev;
    ownsDeviceMemory_ = o.ownsDeviceMemory_;
    o.d_curr = o.d_prev = nullptr;
    o.ownsDeviceMemory_ = false;
    return *this;
}

ScalarField::~ScalarField() { freeDevice(); }

void ScalarField::fill(double v) {
    fillCurr(v);
    fillPrev(v);
}
void ScalarField::fillCurr(double v) { std::fill(curr.begin(), curr.end(), v); }
void ScalarField::fillPrev(double v) { std::fill(prev.begin(), prev.end(), v); }

void ScalarField::advanceTimeLevelCPU() { std::copy(curr.begin(), curr.end(), prev.begin()); }

void ScalarField::advanceTimeLevelGPU() {
    if (!deviceAllocated()) throw std::runtime_error("ScalarField::advanceTimeLevelGPU: device not allocated");
    PHIX_HIP_CHECK(hipMemcpy(d_prev, d_curr, storedSize * sizeof(double), hipMemcpyDeviceToDevice));
}

void ScalarField::allocDevice() {
    if (deviceAllocated()) return;
    const std::size_t bytes = storedSize * sizeof(double);
    PHIX_HIP_CHECK(hipMalloc((void**)&d_curr, bytes));
    PHIX_HIP_CHECK(hipMalloc((void**)&d_prev, bytes));
    PHIX_HIP_CHECK(hipMemset(d_curr, 0, bytes));
    PHIX_HIP_CHECK(hipMemset(d_prev, 0, bytes));
}

void ScalarField::freeDevice() {
    if (ownsDeviceMemory_) {
        if (d_curr) (void)hipFree(d_curr);
        if (d_prev) (void)hipFree(d_prev);
    }
    d_curr = d_prev = nullptr;
}

static void requireDev(const ScalarField& f, const char* who) {
    if (!f.deviceAllocated()) throw std::runtime_error(std::string("ScalarField::") + who + ": device not allocated");
}

void ScalarField::uploadCurrToDevice() const {
    requireDev(*this, "uploadCurrToDevice");
    PHIX_HIP_CHECK(hipMemcpy(d_curr, curr.data(), storedSize * sizeof(double), hipMemcpyHostToDevice));
}
void ScalarField::uploadPrevToDevice() const {
    requireDev(*this, "uploadPrevToDevice");
    PHIX_HIP_CHECK(hipMemcpy(d_prev, prev.data(), storedSize * sizeof(double), hipMemcpyHostToDevice));
}
void ScalarField::uploadAllToDevice() const {
    uploadCurrToDevice();
    uploadPrevToDevice();
}
void ScalarField::downloadCurrFromDevice() {
    requireDev(*this, "downloadCurrFromDevice");
    PHIX_HIP_CHECK(hipMemcpy(curr.data(), d_curr, storedSize * sizeof(double), hipMemcpyDeviceToHost));
}
void ScalarField::downloadPrevFromDevice() {
    requireDev(*this, "downloadPrevFromDevice");
    PHIX_HIP_CHECK(hipMemcpy(prev.data(), d_prev, storedSize * sizeof(double), hipMemcpyDeviceToHost));
}
void ScalarField::downloadAllFromDevice() {
    downloadCurrFromDevice();
    downloadPrevFromDevice();
}

// ------------------------------------------------------------------- IO ----
namespace {

std::size_t physCount(const Mesh& m) { return (std::size_t)m.n[0] * m.n[1] * m.n[2]; }

template <typename Fn>
void forPhys(const ScalarField& f, Fn fn) {
    for (int k = 0; k < f.mesh.n[2]; ++k)
        for (int j = 0; j < f.mesh.n[1]; ++j)
            for (int i = 0; i < f.mesh.n[0]; ++i) fn(i, j, k, (std::size_t)f.index(i, j, k));
}

struct Header {
    std::string name;
    int nx = 1, ny = 1, nz = 1, ghost = 1, ncomp = 1;
};

// Reads "key value" lines up to the "---" terminator.
Header readHeader(std::istream& is, const std::string& path, const char* magic) {
    Header h;
    std::string line;
    if (!std::getline(is, line) || line.find(magic) == std::string::npos)
        throw std::runtime_error("IO: " + path + " is not a " + magic + " file");
    bool done = false;
    while (std::getline(is, line)) {
        if (line == "---" || line == "---\r") {
            done = true;
            break;
        }
        std::istringstream ss(line);
        std::string tok;
        while (ss >> tok) {
            if (tok == "name") ss >> h.name;
            else if (tok == "nx") ss >> h.nx;
            else if (tok == "ny") ss >> h.ny;
            else if (tok == "nz") ss >> h.nz;
            else if (tok == "ghost") ss >> h.ghost;
            else if (tok == "nComponents") ss >> h.ncomp;
        }
    }
    if (!done) throw std::runtime_error("IO: missing '---' header terminator in " + path);
    return h;
}

// VTK XML StructuredGrid of the physical cells: corner points (n+1 per
// axis, from origin + i*d), cell-centred data with one array of `ncomp`
// components (the layout of the reference's VTS writer, FieldIO.cpp:263-312).
void writeVts(const std::string& path, const Mesh& mesh, const std::string& name,
              const std::vector<const ScalarField*>& comps) {
    std::ofstream os(path);
    if (!os) throw std::runtime_error("IO::writeField: cannot open file: " + path);
    const int nx = mesh.n[0], ny = mesh.n[1], nz = mesh.n[2];
    const std::string ext = "0 " + std::to_string(nx) + " 0 " + std::to_string(ny) + " 0 " + std::to_string(nz);
    os << "<?xml version=\"1.0\"?>\n<VTKFile type=\"StructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\">\n"
       << "  <StructuredGrid WholeExtent=\"" << ext << "\">\n    <Piece Extent=\"" << ext << "\">\n"
       << "      <Points>\n        <DataArray type=\"Float64\" NumberOfComponents=\"3\" format=\"ascii\">\n";
    os << std::scientific << std::setprecision(12);
    for (int k = 0; k <= nz; ++k)
        for (int j = 0; j <= ny; ++j)
            for (int i = 0; i <= nx; ++i)
                os << "          " << mesh.origin[0] + i * mesh.d[0] << " " << mesh.origin[1] + j * mesh.d[1] << " "
                   << mesh.origin[2] + k * mesh.d[2] << "\n";
    os << "        </DataArray>\n      </Points>\n      <CellData " << (comps.size() == 1 ? "Scalars" : "Vectors")
       << "=\"" << name << "\">\n        <DataArray type=\"Float64\" Name=\"" << name << "\"";
    if (comps.size() > 1) os << " NumberOfComponents=\"" << comps.size() << "\"";
    os << " format=\"ascii\">\n";
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                os << "         ";
                for (const ScalarField* f : comps) os << " " << f->curr[(std::size_t)f->index(i, j, k)];
                os << "\n";
            }
    os << "        </DataArray>\n      </CellData>\n    </Piece>\n  </StructuredGrid>\n</VTKFile>\n";
    if (!os) throw std::runtime_error("IO::writeField: write failed: " + path);
}

void checkDims(const Header& h, const Mesh& m, const std::string& path) {
    if (h.nx != m.n[0] || h.ny != m.n[1] || h.nz != m.n[2])
        throw std::runtime_error("IO: mesh dimensions in file (" + std::to_string(h.nx) + "x" + std::to_string(h.ny) +
                                 "x" + std::to_string(h.nz) + ") do not match provided mesh (" +
                                 std::to_string(m.n[0]) + "x" + std::to_string(m.n[1]) + "x" +
                                 std::to_string(m.n[2]) + ") in " + path);
}

}  // namespace

void ScalarField::write(const std::string& path, FieldFormat fmt) const {
    if (curr.size() != storedSize) throw std::runtime_error("ScalarField::write: field has no host storage");
    if (fmt == FieldFormat::BINARY) {
        std::ofstream os(path, std::ios::binary);
        if (!os) throw std::runtime_error("IO::writeField: cannot open file: " + path);
        os << "# PhiX ScalarField\nname    " << name << "\nnx " << mesh.n[0] << "  ny " << mesh.n[1] << "  nz "
           << mesh.n[2] << "\nghost   " << ghost << "\n---\n";
        std::vector<double> buf;
        buf.reserve(physCount(mesh));
        forPhys(*this, [&](int, int, int, std::size_t q) { buf.push_back(curr[q]); });
        os.write(reinterpret_cast<const char*>(buf.data()), (std::streamsize)(buf.size() * sizeof(double)));
        if (!os) throw std::runtime_error("IO::writeField: write failed: " + path);
    } else if (fmt == FieldFormat::DAT) {
        std::ofstream os(path);
        if (!os) throw std::runtime_error("IO::writeField: cannot open file: " + path);
        // header lines as the reference writes them (FieldIO.cpp:244-248);
        // values in scientific notation with 17 significant digits (lossless)
        os << "# PhiX ScalarField - DAT\n# name: " << name << "\n# nx " << mesh.n[0] << "  ny " << mesh.n[1]
           << "  nz " << mesh.n[2] << "\n# x y z value\n";
        os << std::scientific << std::setprecision(16);
        forPhys(*this, [&](int i, int j, int k, std::size_t q) {
            os << mesh.coord(0, i) << " " << mesh.coord(1, j) << " " << mesh.coord(2, k) << " " << curr[q] << "\n";
        });
    } else {
        writeVts(path, mesh, name, {this});
    }
}

ScalarField ScalarField::readFromFile(const Mesh& m, const std::string& path, int g) {
    std::ifstream is(path, std::ios::binary);
    if (!is) throw std::runtime_error("IO::readScalarField: cannot open file: " + path);
    Header h = readHeader(is, path, "# PhiX ScalarField");
    checkDims(h, m, path);
    ScalarField f(m, h.name, g);
    std::vector<double> buf(physCount(m));
    is.read(reinterpret_cast<char*>(buf.data()), (std::streamsize)(buf.size() * sizeof(double)));
    if (!is) throw std::runtime_error("IO::readScalarField: unexpected end of binary data in " + path);
    std::size_t t = 0;
    forPhys(f, [&](int, int, int, std::size_t q) { f.curr[q] = buf[t++]; });
    return f;
}

void ScalarField::print() const {
    std::cout << "=== ScalarField ===\n  name       : " << name << "\n  ghost      : " << ghost
              << "\n  storedDims : " << storedDims[0] << " x " << storedDims[1] << " x " << storedDims[2] << "  ("
              << storedSize << " cells)\n  device     : " << (deviceAllocated() ? "allocated" : "not allocated")
              << "\n";
    if (curr.size() != storedSize) return;
    double lo = std::numeric_limits<double>::max(), hi = std::numeric_limits<double>::lowest(), s = 0.0;
    std::size_t cnt = 0;
    forPhys(*this, [&](int, int, int, std::size_t q) {
        lo = std::min(lo, curr[q]);
        hi = std::max(hi, curr[q]);
        s += curr[q];
        ++cnt;
    });
    if (cnt)
        std::cout << "  curr  min  : " << lo << "\n  curr  max  : " << hi << "\n  curr  mean : " << s / (double)cnt
                  << "\n";
}

// ----------------------------------------------------------- VectorField ----
std::string VectorField::componentName(const std::string& base, int c, int n) {
    if (n == 3) {
        const char* xyz[3] = {"_x", "_y", "_z"};
        return base + xyz[c];
    }
    return base + "_" + std::to_string(c);
}

VectorField::VectorField(const Mesh& m, const std::string& nm, int nComp, int g) : mesh(m), ghost(g), name(nm) {
    if (nComp < 1) throw std::invalid_argument("VectorField: nComponents must be >= 1");
    components_.reserve(nComp);
    for (int c = 0; c < nComp; ++c) components_.emplace_back(m, componentName(nm, c, nComp), g);
}

VectorField::VectorField(VectorField&& o) noexcept
    : mesh(o.mesh), ghost(o.ghost), name(std::move(o.name)), components_(std::move(o.components_)) {}

VectorField& VectorField::operator=(VectorField&& o) noexcept {
    if (this != &o) {
        ghost = o.ghost;
        name = std::move(o.name);
        components_ = std::move(o.components_);
    }
    return *this;
}

ScalarField& VectorField::operator[](int c) {
    if (c < 0 || c >= nComponents()) throw std::out_of_range("VectorField: component index out of range");
    return components_[c];
}
const ScalarField& VectorField::operator[](int c) const {
    if (c < 0 || c >= nComponents()) throw std::out_of_range("VectorField: component index out of range");
    return components_[c];
}

void VectorField::fill(double v) { for (auto& f : components_) f.fill(v); }
void VectorField::fillCurr(double v) { for (auto& f : components_) f.fillCurr(v); }
void VectorField::fillPrev(double v) { for (auto& f : components_) f.fillPrev(v); }
void VectorField::advanceTimeLevelCPU() { for (auto& f : components_) f.advanceTimeLevelCPU(); }
void VectorField::advanceTimeLevelGPU() { for (auto& f : components_) f.advanceTimeLevelGPU(); }
bool VectorField::deviceAllocated() const {
    for (auto& f : components_)
        if (!f.deviceAllocated()) return false;
    return true;
}
void VectorField::allocDevice() { for (auto& f : components_) f.allocDevice(); }
void VectorField::freeDevice() { for (auto& f : components_) f.freeDevice(); }
void VectorField::uploadCurrToDevice() const { for (auto& f : components_) f.uploadCurrToDevice(); }
void VectorField::uploadPrevToDevice() const { for (auto& f : components_) f.uploadPrevToDevice(); }
void VectorField::uploadAllToDevice() const { for (auto& f : components_) f.uploadAllToDevice(); }
void VectorField::downloadCurrFromDevice() { for (auto& f : components_) f.downloadCurrFromDevice(); }
void VectorField::downloadPrevFromDevice() { for (auto& f : components_) f.downloadPrevFromDevice(); }
void VectorField::downloadAllFromDevice() { for (auto& f : components_) f.downloadAllFromDevice(); }

void VectorField::write(const std::string& path, FieldFormat fmt) const {
    for (const auto& f : components_)
        if (f.curr.size() != f.storedSize) throw std::runtime_error("VectorField::write: field has no host storage");
    if (fmt == FieldFormat::VTS) {
        std::vector<const ScalarField*> comps;
        for (const auto& f : components_) comps.push_back(&f);
        writeVts(path, mesh, name, comps);
        return;
    }
    if (fmt == FieldFormat::DAT) {
        std::ofstream os(path);
        if (!os) throw std::runtime_error("IO::writeField: cannot open file: " + path);
        // header lines as the reference writes them (FieldIO.cpp:358-363)
        os << "# PhiX VectorField - DAT\n# name: " << name << "  nComponents: " << nComponents() << "\n# nx "
           << mesh.n[0] << "  ny " << mesh.n[1] << "  nz " << mesh.n[2] << "\n# x y z";
        for (int c = 0; c < nComponents(); ++c) os << " v" << c;
        os << "\n" << std::scientific << std::setprecision(16);
        forPhys(components_[0], [&](int i, int j, int k, std::size_t q) {
            os << mesh.coord(0, i) << " " << mesh.coord(1, j) << " " << mesh.coord(2, k);
            for (const auto& f : components_) os << " " << f.curr[q];
            os << "\n";
        });
        return;
    }
    std::ofstream os(path, std::ios::binary);
    if (!os) throw std::runtime_error("IO::writeField: cannot open file: " + path);
    os << "# PhiX VectorField\nname         " << name << "\nnComponents  " << nComponents() << "\nnx " << mesh.n[0]
       << "  ny " << mesh.n[1] << "  nz " << mesh.n[2] << "\nghost        " << ghost << "\n---\n";
    std::vector<double> buf;
    for (const auto& f : components_) {
        buf.clear();
        forPhys(f, [&](int, int, int, std::size_t q) { buf.push_back(f.curr[q]); });
        os.write(reinterpret_cast<const char*>(buf.data()), (std::streamsize)(buf.size() * sizeof(double)));
    }
}

VectorField VectorField::readFromFile(const Mesh& m, const std::string& path, int g) {
    std::ifstream is(path, std::ios::binary);
    if (!is) throw std::runtime_error("IO::readVectorField: cannot open file: " + path);
    Header h = readHeader(is, path, "# PhiX VectorField");
    checkDims(h, m, path);
    VectorField vf(m, h.name, h.ncomp, g);
    std::vector<double> buf(physCount(m));
    for (int c = 0; c < h.ncomp; ++c) {
        is.read(reinterpret_cast<char*>(buf.data()), (std::streamsize)(buf.size() * sizeof(double)));
        if (!is) throw std::runtime_error("IO::readVectorField: unexpected end of binary data in " + path);
        std::size_t t = 0;
        forPhys(vf[c], [&](int, int, int, std::size_t q) { vf[c].curr[q] = buf[t++]; });
    }
    return vf;
}

void VectorField::print() const {
    for (const auto& f : components_) f.print();
}

}  // namespace PhiX
