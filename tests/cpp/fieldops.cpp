// Known-answer tests of pw() and the field operators through the drop-in C++
// API -- the cases and expected values of the reference's
// develop/fieldOperation/test_fieldops.cu:118-291 (4x4 mesh, ghost 1, f1=2,
// f2=3, f3=5, tolerance 1e-12), written against the drop-in headers.
// Usage: fieldops cpu|gpu   (gpu also checks the device path)
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>

#include "equation/Equation.h"
#include "equation/VectorEquation.h"
#include "mesh/Mesh.h"

using namespace PhiX;

static int g_fail = 0, g_total = 0;

static bool uniform(const ScalarField& f, double want) {
    for (int j = 0; j < f.mesh.n[1]; ++j)
        for (int i = 0; i < f.mesh.n[0]; ++i)
            if (std::fabs(f.curr[f.index(i, j)] - want) > 1e-12) {
                std::printf("    mismatch at (%d,%d): got %.17g want %.17g\n", i, j, f.curr[f.index(i, j)], want);
                return false;
            }
    return true;
}

static void report(const std::string& what, bool ok) {
    ++g_total;
    if (!ok) ++g_fail;
    std::printf("  [%s] %s\n", ok ? "PASS" : "FAIL", what.c_str());
}

static bool g_gpu = false;

static void check(const std::string& name, ScalarField& primary, const Term& t, ScalarField& rhs, double want) {
    Equation eq(primary, name);
    eq.setRHS(t);
    eq.computeRHSCPU(rhs);
    report(name + " [CPU]", uniform(rhs, want));
    if (g_gpu) {
        eq.computeRHS(rhs);
        rhs.downloadCurrFromDevice();
        report(name + " [GPU]", uniform(rhs, want));
    }
}

int main(int argc, char** argv) {
    g_gpu = argc > 1 && std::strcmp(argv[1], "gpu") == 0;
    Mesh mesh = Mesh::makeUniform2D(CoordSys::CARTESIAN, 4, 1.0, 0.0, 4, 1.0, 0.0);
    ScalarField f1(mesh, "f1", 1), f2(mesh, "f2", 1), f3(mesh, "f3", 1), rhs(mesh, "rhs", 1);
    f1.fillCurr(2.0);
    f2.fillCurr(3.0);
    f3.fillCurr(5.0);
    if (g_gpu) {
        f1.allocDevice(); f1.uploadAllToDevice();
        f2.allocDevice(); f2.uploadAllToDevice();
        f3.allocDevice(); f3.uploadAllToDevice();
        rhs.allocDevice();
    }
    check("pw(f1, v^2)", f1, pw(f1, PHIX_FN(double v) { return v * v; }), rhs, 4.0);
    check("pw(f1, v^2, coeff=3)", f1, pw(f1, PHIX_FN(double v) { return v * v; }, 3.0), rhs, 12.0);
    check("pw(f1, f2, a*b)", f1, pw(f1, f2, PHIX_FN(double a, double b) { return a * b; }), rhs, 6.0);
    check("pw(f1, f2, a+b)", f1, pw(f1, f2, PHIX_FN(double a, double b) { return a + b; }), rhs, 5.0);
    check("pw(f1, f2, a*b, coeff=2)", f1, pw(f1, f2, PHIX_FN(double a, double b) { return a * b; }, 2.0), rhs, 12.0);
    check("pw(f1, f2, f3, a+b+c)", f1, pw(f1, f2, f3, PHIX_FN(double a, double b, double c) { return a + b + c; }), rhs,
          10.0);
    check("pw(f1, f2, f3, a*b*c)", f1, pw(f1, f2, f3, PHIX_FN(double a, double b, double c) { return a * b * c; }), rhs,
          30.0);
    check("pw(f1, f2, f3, a+b+c, coeff=0.5)", f1,
          pw(f1, f2, f3, PHIX_FN(double a, double b, double c) { return a + b + c; }, 0.5), rhs, 5.0);
    check("f1 * f2", f1, f1 * f2, rhs, 6.0);
    check("f1 + f2", f1, f1 + f2, rhs, 5.0);
    check("f1 - f2", f1, f1 - f2, rhs, -1.0);
    check("3.0 * f1", f1, 3.0 * f1, rhs, 6.0);
    check("f1 * 4.0", f1, f1 * 4.0, rhs, 8.0);
    check("f1 + 1.0", f1, f1 + 1.0, rhs, 3.0);
    check("1.0 + f1", f1, 1.0 + f1, rhs, 3.0);
    check("f1 - 1.0", f1, f1 - 1.0, rhs, 1.0);
    check("10.0 - f1", f1, 10.0 - f1, rhs, 8.0);
    check("-f1", f1, -f1, rhs, -2.0);
    {
        RHSExpr e(f1 * f2);
        e += pw(f3, PHIX_FN(double v) { return v; });
        Equation eq(f1, "compose");
        eq.setRHS(e);
        eq.computeRHSCPU(rhs);
        report("(f1*f2) + pw(f3, id) [CPU]", uniform(rhs, 11.0));
        if (g_gpu) {
            eq.computeRHS(rhs);
            rhs.downloadCurrFromDevice();
            report("(f1*f2) + pw(f3, id) [GPU]", uniform(rhs, 11.0));
        }
    }
    // composite products (Term * Term materialises both operands)
    check("(f1*f2) * (f3 + 1.0)", f1, (f1 * f2) * (f3 + 1.0), rhs, 36.0);
    check("(f1*f2) * f3", f1, (f1 * f2) * f3, rhs, 30.0);
    {
        VectorField vf(mesh, "vf", 2, 1), vr(mesh, "vr", 2, 1);
        vf.fillCurr(2.0);
        VectorEquation ve(vf, "vf_sf");
        ve.setRHS(pw(vf, f2, PHIX_FN(double a, double b) { return a * b; }));
        ve.computeRHSCPU(vr);
        report("pw(vf, sf, a*b) [CPU]", uniform(vr[0], 6.0) && uniform(vr[1], 6.0));
        ve.setRHS(pw(vf, f2, PHIX_FN(double a, double b) { return a * b; }, 0.5));
        ve.computeRHSCPU(vr);
        report("pw(vf, sf, a*b, coeff=0.5) [CPU]", uniform(vr[0], 3.0) && uniform(vr[1], 3.0));
    }
    {
        VectorField v1(mesh, "v1", 2, 1), v2(mesh, "v2", 2, 1), vr(mesh, "vr2", 2, 1);
        v1.fillCurr(2.0);
        v2.fillCurr(4.0);
        VectorEquation ve(v1, "vv");
        ve.setRHS(pw(v1, v2, PHIX_FN(double a, double b) { return a + b; }));
        ve.computeRHSCPU(vr);
        report("pw(vf1, vf2, a+b) [CPU]", uniform(vr[0], 6.0) && uniform(vr[1], 6.0));
        ve.setRHS(pw(v1, v2, PHIX_FN(double a, double b) { return a * b; }));
        ve.computeRHSCPU(vr);
        report("pw(vf1, vf2, a*b) [CPU]", uniform(vr[0], 8.0) && uniform(vr[1], 8.0));
    }
    // planner probes: the double-well and Allen-Cahn forms are recognised
    {
        const double rho = 5.0, ca = 0.3, cb = 0.7;
        Term t = pw(f1, PHIX_FN(double c) { return 2.0 * rho * (c - ca) * (c - cb) * (2.0 * c - ca - cb); });
        report("probe double-well (5, 0.3, 0.7)", t.sym.doublewell && t.sym.dw[0] == 5.0 && t.sym.dw[1] == 0.3 &&
                                                      t.sym.dw[2] == 0.7);
        Term u = pw(f1, PHIX_FN(double c) { return c * c * c; });
        report("probe rejects c^3", !u.sym.doublewell && !u.sym.ac_well);
        Term w = pw(f1, PHIX_FN(double e) { return e - e * e * e; });
        report("probe Allen-Cahn well", w.sym.ac_well && !w.sym.doublewell);
        Term x = pw(f1, f2, PHIX_FN(double a, double b) { return a * b * b; });
        report("probe a*b*b", x.sym.abb);
    }
    // error behaviour
    {
        Mesh m2 = Mesh::makeUniform2D(CoordSys::CARTESIAN, 5, 1.0, 0.0, 4, 1.0, 0.0);
        ScalarField g(m2, "g", 1);
        bool threw = false;
        try {
            (void)pw(f1, g, PHIX_FN(double a, double b) { return a + b; });
        } catch (const std::invalid_argument&) {
            threw = true;
        }
        report("pw shape mismatch throws invalid_argument", threw);
        Equation eq(f1, "empty");
        threw = false;
        try {
            eq.computeRHSCPU(rhs);
        } catch (const std::runtime_error&) {
            threw = true;
        }
        report("computeRHSCPU before setRHS throws", threw);
        threw = false;
        try {
            (void)grad(f1, 2);
        } catch (const std::invalid_argument&) {
            threw = true;
        }
        report("grad axis out of range throws", threw);
        threw = false;
        try {
            Mesh bad(2, CoordSys::CARTESIAN, 4, 1.0, 0.0, 0, 1.0, 0.0, 1, 1.0, 0.0);
        } catch (const std::invalid_argument&) {
            threw = true;
        }
        report("Mesh validate throws", threw);
    }
    std::printf("%d/%d passed\n", g_total - g_fail, g_total);
    return g_fail == 0 ? 0 : 1;
}
