// Composite expression forms of the drop-in DSL on non-uniform fields: the
// right-hand sides are evaluated by the host path (computeRHSCPU) and, with
// "gpu", by the device path (computeRHS), and written as .field files for the
// comparison with a numpy restatement (tests/test_cpp_api_cpu.py,
// tests/test_gpu_cpp_api.py). Forms (include/equation/FieldOps.inl:136-250,
// src/equation/Equation.cu:655-743 in the reference):
//   gradx_sq   grad(pw(a, v^2), 0, bcs)           composite gradient of an expression
//   grady_ab   grad(a * b, 1, bcs)
//   lap_ab     lap(a * b, bcs)                     composite Laplacian
//   divgrad_ab div(grad(a * b, bcs), bcs)          divergence of a composite vector expression
//   ab_lapb    (a * b) * lap(b)                    Term * Term
//   mixed      2 grad(a, 0) + div(v) - a * b       field gradient, VectorField divergence, sum
// Usage: composite cpu|gpu <outdir> periodic|noflux
// Mesh 24 x 20, dx = 0.5, dy = 0.25; a = 1.5 + sin(x) cos(2y), b = cos(0.7x - 1.3y);
// the ghost cells of a, b and v are BC-filled before evaluation.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

#include "boundary/NoFluxBC.h"
#include "boundary/PeriodicBC.h"
#include "equation/Equation.h"
#include "equation/VectorEquation.h"
#include "mesh/Mesh.h"

using namespace PhiX;

int main(int argc, char** argv) {
    if (argc < 4) {
        std::fprintf(stderr, "usage: composite cpu|gpu outdir periodic|noflux\n");
        return 2;
    }
    const bool gpu = std::strcmp(argv[1], "gpu") == 0;
    const std::string out = argv[2], mode = argv[3];
    try {
        Mesh mesh = Mesh::makeUniform2D(CoordSys::CARTESIAN, 24, 0.5, 0.0, 20, 0.25, 0.0);
        ScalarField a(mesh, "a", 1), b(mesh, "b", 1), rhs(mesh, "rhs", 1);
        VectorField v(mesh, "v", 2, 1);
        for (int j = 0; j < 20; ++j)
            for (int i = 0; i < 24; ++i) {
                const double x = mesh.coord(0, i), y = mesh.coord(1, j);
                a.curr[a.index(i, j)] = 1.5 + std::sin(x) * std::cos(2.0 * y);
                b.curr[b.index(i, j)] = std::cos(0.7 * x - 1.3 * y);
                v[0].curr[v[0].index(i, j)] = x * std::sin(y);
                v[1].curr[v[1].index(i, j)] = std::exp(-0.1 * x) * y;
            }
        std::vector<std::unique_ptr<BoundaryCondition>> own;
        if (mode == "periodic") {
            own.push_back(std::make_unique<PeriodicBC>(Axis::X));
            own.push_back(std::make_unique<PeriodicBC>(Axis::Y));
        } else {
            own.push_back(std::make_unique<NoFluxBC>(Axis::X));
            own.push_back(std::make_unique<NoFluxBC>(Axis::Y));
        }
        std::vector<BoundaryCondition*> bcs;
        for (auto& p : own) bcs.push_back(p.get());
        for (auto* bc : bcs) {
            bc->applyOnCPU(a);
            bc->applyOnCPU(b);
            bc->applyOnCPU(v);
        }
        if (gpu) {
            for (ScalarField* f : {&a, &b, &rhs, &v[0], &v[1]}) {
                f->allocDevice();
                f->uploadAllToDevice();
            }
        }
        struct Case {
            const char* name;
            std::function<RHSExpr()> make;
        };
        const std::vector<Case> cases = {
            {"gradx_sq", [&] { return RHSExpr(grad(pw(a, PHIX_FN(double u) { return u * u; }), 0, bcs)); }},
            {"grady_ab", [&] { return RHSExpr(grad(a * b, 1, bcs)); }},
            {"lap_ab", [&] { return RHSExpr(lap(a * b, bcs)); }},
            {"divgrad_ab", [&] { return div(grad(a * b, bcs), bcs); }},
            {"ab_lapb", [&] { return RHSExpr((a * b) * lap(b)); }},
            {"mixed", [&] {
                 RHSExpr e(grad(a, 0, 2.0));
                 e += div(v);
                 e += -1.0 * (a * b);
                 return e;
             }},
        };
        for (const auto& c : cases) {
            Equation eq(a, c.name);
            eq.setRHS(c.make());
            if (gpu) {
                eq.computeRHS(rhs);
                rhs.downloadCurrFromDevice();
            } else {
                eq.computeRHSCPU(rhs);
            }
            rhs.write(out + "/" + c.name + "_" + mode + "_" + (gpu ? "gpu" : "cpu") + ".field");
            std::printf("%s done\n", c.name);
        }
    } catch (const std::exception& e) {
        std::fprintf(stderr, "composite: %s\n", e.what());
        return 1;
    }
    return 0;
}
