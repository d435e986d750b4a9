// Host-only checks of the configuration and output layer of the drop-in C++
// API (no device calls):
//   * IO::Json / ConfigFile: JSONC parsing (comments, escapes, numbers, the
//     reference's own sample settings files), conversions, errors with
//     line:column, dump() round trip;
//   * buildBCs: the reference's rules (src/boundary/BCFactory.cpp:13-74);
//   * OutputWriter: cadence and the files writeFields produces;
//   * FieldIO: BINARY and DAT round trips (bit-exact), mesh-mismatch errors,
//     resolveStartStep's nearest-file rule (src/IO/FieldIO.cpp:750-782) and
//     initField's paths.
// Usage: config_io <scratch dir> <settings.jsonc> [<settings.jsonc> ...]
// The scratch directory becomes the working directory (output/ and
// settings/initial_field/ are created there). Prints one line per check and
// "FAIL" for each failure; exits 1 if any failed.
#include <unistd.h>

#include <cmath>
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <functional>
#include <string>
#include <typeinfo>

#include "IO/ConfigFile.h"
#include "IO/FieldIO.h"
#include "IO/OutputWriter.h"
#include "boundary/BCFactory.h"
#include "boundary/NoFluxBC.h"
#include "boundary/PeriodicBC.h"
#include "field/ScalarField.h"
#include "field/VectorField.h"

using namespace PhiX;
using IO::Json;

static int g_fail = 0, g_total = 0;

static void report(const std::string& what, bool ok) {
    ++g_total;
    if (!ok) ++g_fail;
    std::printf("%s  %s\n", ok ? "ok  " : "FAIL", what.c_str());
}

// Runs f and reports whether it threw an exception of type E whose message
// contains `needle`.
template <class E>
static void expectThrow(const std::string& what, const std::function<void()>& f, const std::string& needle = "") {
    try {
        f();
        report(what + " (no exception)", false);
    } catch (const E& e) {
        const bool ok = needle.empty() || std::string(e.what()).find(needle) != std::string::npos;
        report(what + (ok ? "" : std::string(" (message: ") + e.what() + ")"), ok);
    } catch (const std::exception& e) {
        report(what + " (wrong exception " + typeid(e).name() + ": " + e.what() + ")", false);
    }
}

static void writeText(const std::string& path, const std::string& text) {
    std::ofstream os(path);
    os << text;
}

static void jsonChecks() {
    const std::string text = R"({
        // comment line
        "mesh": { "nx": 200, "dx": 1.0, "x0": -2.5e-1 },   // trailing comment
        "url": "http://example.org//path",                 // '//' inside a string
        "esc": "a\"b\\c\/d\n\u00e9\ud83d\ude00",
        "list": [1, -2, 3.5, true, false, null, [], {}],
        "big": 9007199254740993,
        "exp": 1E+2,
        "dup": 1, "dup": 2
    })";
    const Json j = Json::parse(text);
    const int nx = j["mesh"]["nx"];
    const double dx = j["mesh"]["dx"], x0 = j["mesh"]["x0"];
    report("json: nested numbers convert (int and double)", nx == 200 && dx == 1.0 && x0 == -0.25);
    report("json: an integer converts to double", j["mesh"]["nx"].get<double>() == 200.0);
    report("json: '//' inside a string is kept", j["url"].get<std::string>() == "http://example.org//path");
    report("json: escapes and surrogate pairs decode to UTF-8",
           j["esc"].get<std::string>() == "a\"b\\c/d\n\xc3\xa9\xf0\x9f\x98\x80");
    const Json& l = j["list"];
    report("json: array elements and types",
           l.size() == 8 && l[0].get<int>() == 1 && l[1].get<int>() == -2 && l[2].get<double>() == 3.5 &&
               l[3].get<bool>() && !l[4].get<bool>() && l[5].is_null() && l[6].is_array() && l[6].size() == 0 &&
               l[7].is_object() && l[7].size() == 0);
    report("json: 64-bit integers stay exact", j["big"].get<long long>() == 9007199254740993LL);
    report("json: exponent form is a non-integral number", j["exp"].get<double>() == 100.0 && !j["exp"].is_number_integer());
    report("json: a repeated key keeps its last value", j["dup"].get<int>() == 2 && j.size() == 7);
    report("json: value() falls back to the default",
           j.value("missing", 7) == 7 && j.value("dup", 7) == 2 && j.value("nokey", "FP64") == std::string("FP64"));
    report("json: contains", j.contains("mesh") && !j.contains("nope") && !j["mesh"]["nx"].contains("x"));
    const Json r = Json::parse(j.dump());
    report("json: dump() round trip", r.dump() == j.dump() && r["esc"].get<std::string>() == j["esc"].get<std::string>() &&
                                          r["mesh"]["x0"].get<double>() == -0.25);
    const double third = Json::parse(Json::parse("[0.1, 0.3333333333333333]").dump())[1].get<double>();
    report("json: dump() keeps doubles exact", third == 0.3333333333333333);

    expectThrow<std::out_of_range>("json: missing key -> std::out_of_range", [&] { (void)j.at("nope"); }, "nope");
    expectThrow<std::out_of_range>("json: index past the end -> std::out_of_range", [&] { (void)l.at(8); });
    expectThrow<std::runtime_error>("json: string read as number -> type error",
                                    [&] { (void)j["url"].get<double>(); }, "expected a number");
    expectThrow<std::runtime_error>("json: number read as string -> type error",
                                    [&] { (void)j["exp"].get<std::string>(); }, "expected a string");
    const std::pair<const char*, const char*> bad[] = {
        {"{\"a\": 1,}", "line 1, column 9"},                // trailing comma: a key is expected
        {"{\"a\" 1}", "expected ':'"},
        {"{\n  \"a\": tru\n}", "line 2, column 8"},       // bad literal
        {"{\"a\": \"unterminated}", "unterminated string"},
        {"[1, 2", "expected ',' or ']'"},
        {"{\"a\": 01}", "line 1, column 8"},                // a leading zero ends the number
        {"{} x", "unexpected characters after the value"},
        {"{\"a\": 1.}", "digits expected after '.'"},
        {"\"\\ud800\"", "unpaired surrogate"},
        {"", "unexpected end of input"},
    };
    for (const auto& b : bad)
        expectThrow<std::runtime_error>(std::string("json: rejects ") + b.first,
                                        [&] { (void)Json::parse(b.first); }, b.second);
    std::string deep(600, '[');
    expectThrow<std::runtime_error>("json: rejects nesting beyond 512 levels", [&] { (void)Json::parse(deep); },
                                    "nesting");
}

static void settingsChecks(int argc, char** argv) {
    for (int a = 2; a < argc; ++a) {
        const IO::ConfigFile cfg(argv[a]);
        const bool ok = cfg.has("mesh") && cfg.has("initialize") && cfg.has("boundary_conditions") &&
                        cfg.has("constants") && cfg.has("output") && cfg["mesh"]["nx"].get<int>() == 200 &&
                        cfg["initialize"]["start_from"].get<std::string>() == "initial_field" &&
                        cfg["initialize"]["dt"].get<double>() == 0.001;
        report(std::string("settings: ") + argv[a] + " parses with the reference's keys", ok);
        const BCSet bcs = buildBCs(cfg["boundary_conditions"]);
        report(std::string("settings: ") + argv[a] + " builds two periodic BCs", bcs.ptrs.size() == 2 &&
               dynamic_cast<PeriodicBC*>(bcs.ptrs[0]) && dynamic_cast<PeriodicBC*>(bcs.ptrs[1]));
    }
    writeText("bad.jsonc", "{\n  \"mesh\": {\n    \"nx\": 10,,\n  }\n}\n");
    expectThrow<std::runtime_error>("settings: a broken file names its position and path",
                                    [] { IO::ConfigFile c("bad.jsonc"); }, "line 3, column 14");
    expectThrow<std::runtime_error>("settings: a missing file is an error",
                                    [] { IO::ConfigFile c("no/such/settings.jsonc"); }, "cannot open");
}

static void bcChecks() {
    auto build = [](const std::string& text) { return buildBCs(Json::parse(text)); };
    {
        const BCSet s = build(R"({"x_min": "NoFlux", "x_max": "NoFlux", "y_min": "Periodic", "y_max": "Periodic"})");
        auto* lo = dynamic_cast<NoFluxBC*>(s.ptrs.size() == 3 ? s.ptrs[0] : nullptr);
        auto* hi = dynamic_cast<NoFluxBC*>(s.ptrs.size() == 3 ? s.ptrs[1] : nullptr);
        auto* py = dynamic_cast<PeriodicBC*>(s.ptrs.size() == 3 ? s.ptrs[2] : nullptr);
        report("bc: NoFlux per side, Periodic per axis, in x-y order",
               lo && hi && py && lo->side == Side::LOW && hi->side == Side::HIGH && lo->axis == Axis::X &&
                   py->axis == Axis::Y && s.storage.size() == 3);
    }
    report("bc: y and z pairs are optional", build(R"({"x_min": "Periodic", "x_max": "Periodic"})").ptrs.size() == 1);
    report("bc: a lone y_min is ignored",
           build(R"({"x_min": "Periodic", "x_max": "Periodic", "y_min": "NoFlux"})").ptrs.size() == 1);
    report("bc: 3D with z pair", build(R"({"x_min": "Periodic", "x_max": "Periodic", "y_min": "NoFlux",
                                          "y_max": "NoFlux", "z_min": "Periodic", "z_max": "Periodic"})")
                                         .ptrs.size() == 4);
    expectThrow<std::runtime_error>("bc: Periodic on one side only is rejected",
                                    [&] { build(R"({"x_min": "Periodic", "x_max": "NoFlux"})"); }, "both sides");
    expectThrow<std::runtime_error>("bc: an unknown type is rejected",
                                    [&] { build(R"({"x_min": "Dirichlet", "x_max": "NoFlux"})"); }, "Dirichlet");
    expectThrow<std::out_of_range>("bc: x pair is required", [&] { build(R"({"y_min": "NoFlux"})"); });
}

static bool sameField(const ScalarField& a, const ScalarField& b) {
    for (int k = 0; k < a.mesh.n[2]; ++k)
        for (int j = 0; j < a.mesh.n[1]; ++j)
            for (int i = 0; i < a.mesh.n[0]; ++i)
                if (a.curr[a.index(i, j, k)] != b.curr[b.index(i, j, k)]) return false;
    return true;
}

static void fieldIoChecks() {
    namespace fs = std::filesystem;
    const Mesh m = Mesh::makeUniform3D(CoordSys::CARTESIAN, 5, 0.5, 1.0, 4, 0.25, -1.0, 3, 1.0, 0.0);
    ScalarField f(m, "phi", 1);
    for (int k = 0; k < 3; ++k)
        for (int j = 0; j < 4; ++j)
            for (int i = 0; i < 5; ++i) f.curr[f.index(i, j, k)] = std::sin(0.7 * i + 1.3 * j - 0.4 * k) / 3.0 + 1e-300 * i;
    for (const auto fmt : {FieldFormat::BINARY, FieldFormat::DAT}) {
        const std::string path = fmt == FieldFormat::DAT ? "phi.dat" : "phi.field";
        IO::writeField(f, path, fmt);
        ScalarField g(m, "phi", 1);
        g.fill(-7.0);
        IO::readField(g, path);
        const ScalarField h = IO::readScalarField(m, path);
        report("fieldio: " + path + " round trip is bit-exact", sameField(f, g) && sameField(f, h) && h.name == "phi");
        ScalarField ghost2(m, "phi", 2);
        IO::readField(ghost2, path);
        report("fieldio: " + path + " reads into a field with two ghost layers", sameField(f, ghost2));
    }
    {
        std::ifstream is("phi.dat");
        std::string l1, l2, l3;
        std::getline(is, l1);
        std::getline(is, l2);
        std::getline(is, l3);
        report("fieldio: DAT header lines", l1 == "# PhiX ScalarField - DAT" && l2 == "# name: phi" &&
                                                l3 == "# nx 5  ny 4  nz 3");
    }
    const Mesh other = Mesh::makeUniform2D(CoordSys::CARTESIAN, 5, 0.5, 1.0, 4, 0.25, -1.0);
    for (const std::string path : {"phi.field", "phi.dat"}) {
        ScalarField wrong(other, "phi", 1);
        expectThrow<std::runtime_error>("fieldio: mesh mismatch is rejected (" + path + ")",
                                        [&] { IO::readField(wrong, path); }, "do not match");
    }
    expectThrow<std::runtime_error>("fieldio: a missing file is rejected",
                                    [&] { IO::readField(f, "absent.field"); }, "cannot open");

    VectorField v(m, "eta", 3, 1);
    for (int c = 0; c < 3; ++c)
        for (int k = 0; k < 3; ++k)
            for (int j = 0; j < 4; ++j)
                for (int i = 0; i < 5; ++i) v[c].curr[v[c].index(i, j, k)] = std::cos(0.3 * i * (c + 1) + j - k) * 0.1;
    for (const auto fmt : {FieldFormat::BINARY, FieldFormat::DAT}) {
        const std::string path = fmt == FieldFormat::DAT ? "eta.dat" : "eta.vfield";
        IO::writeField(v, path, fmt);
        VectorField w(m, "eta", 3, 1);
        IO::readField(w, path);
        const VectorField x = IO::readVectorField(m, path);
        bool ok = x.nComponents() == 3;
        for (int c = 0; c < 3 && ok; ++c) ok = sameField(v[c], w[c]) && sameField(v[c], x[c]);
        report("fieldio: vector " + path + " round trip is bit-exact", ok);
        VectorField two(m, "eta", 2, 1);
        expectThrow<std::runtime_error>("fieldio: vector component mismatch is rejected (" + path + ")",
                                        [&] { IO::readField(two, path); }, "component");
    }

    // restart resolution: output/c_<step>.field, nearest step wins, ties go
    // to the earlier step
    fs::create_directories("output");
    for (const char* name : {"c_100.field", "c_250.field", "c_1000.field", "c_x.field", "mu_300.field", "c_400.dat"})
        writeText(std::string("output/") + name, "x");
    report("restart: initial_field -> 0", IO::resolveStartStep("initial_field") == 0);
    report("restart: nearest step (240 -> 250)", IO::resolveStartStep("240") == 250);
    report("restart: nearest step (0 -> 100)", IO::resolveStartStep("0") == 100);
    report("restart: tie goes to the earlier step (175 -> 100)", IO::resolveStartStep("175") == 100);
    report("restart: only .field files of the reference field count (320 -> 250)",
           IO::resolveStartStep("320") == 250 && IO::resolveStartStep("320", "mu") == 300);
    report("restart: beyond the last file (5000 -> 1000)", IO::resolveStartStep("5000") == 1000);
    expectThrow<std::invalid_argument>("restart: a non-number is std::invalid_argument",
                                       [] { IO::resolveStartStep("latest"); });
    expectThrow<std::runtime_error>("restart: no files for the field is an error",
                                    [] { IO::resolveStartStep("10", "eta"); }, "no eta_*.field files");
    fs::remove_all("output");

    // initField: step 0 -> settings/initial_field/<name>.field, else output/<name>_<step>.field
    const Mesh m2 = Mesh::makeUniform2D(CoordSys::CARTESIAN, 6, 1.0, 0.0, 4, 1.0, 0.0);
    ScalarField c(m2, "c", 1);
    for (int j = 0; j < 4; ++j)
        for (int i = 0; i < 6; ++i) c.curr[c.index(i, j)] = 0.5 + 0.01 * (i - 2 * j);
    fs::create_directories("settings/initial_field");
    IO::writeField(c, "settings/initial_field/c.field");
    ScalarField cold(m2, "c", 1);
    IO::initField(cold, 0);
    report("initField: cold start reads settings/initial_field/c.field", sameField(c, cold));

    // OutputWriter: cadence, files per format, warm start from what it wrote
    IO::OutputWriter w(Json::parse(R"({"print_interval": 3, "write_interval": 4, "format": "ALL"})"));
    report("output: cadence", w.shouldPrint(6) && !w.shouldPrint(7) && w.shouldWrite(8) && !w.shouldWrite(6) &&
                                  w.shouldPrint(0) && w.shouldWrite(0));
    w.writeFields(c, 8, 0.08);
    w.printProgress(8, 0.08);
    report("output: ALL writes .field, .dat and .vts",
           fs::exists("output/c_8.field") && fs::exists("output/c_8.dat") && fs::exists("output/c_8.vts"));
    std::string vts;
    {
        std::ifstream is("output/c_8.vts");
        vts.assign(std::istreambuf_iterator<char>(is), {});
    }
    report("output: .vts is a StructuredGrid of 7x5x2 points with cell data",
           vts.find("<VTKFile type=\"StructuredGrid\"") != std::string::npos &&
               vts.find("WholeExtent=\"0 6 0 4 0 1\"") != std::string::npos &&
               vts.find("<CellData Scalars=\"c\">") != std::string::npos);
    ScalarField warm(m2, "c", 1);
    report("output: warm start resolves to the written step", IO::resolveStartStep("9") == 8);
    IO::initField(warm, 8);
    report("output: warm start reads back what was written", sameField(c, warm));
    IO::OutputWriter only(Json::parse(R"({"print_interval": 1, "write_interval": 1, "format": "DAT"})"));
    only.writeFields(c, 12, 0.12);
    report("output: DAT writes only .dat", fs::exists("output/c_12.dat") && !fs::exists("output/c_12.field") &&
                                               !fs::exists("output/c_12.vts"));
    expectThrow<std::runtime_error>(
        "output: a zero interval is rejected",
        [] { IO::OutputWriter z(Json::parse(R"({"print_interval": 0, "write_interval": 1, "format": "ALL"})")); },
        "interval");
}

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: config_io <scratch dir> [settings.jsonc ...]\n");
        return 2;
    }
    // settings paths are taken relative to the caller's directory
    for (int a = 2; a < argc; ++a) {
        static std::string keep[64];
        keep[a % 64] = std::filesystem::absolute(argv[a]).string();
        argv[a] = keep[a % 64].data();
    }
    std::filesystem::create_directories(argv[1]);
    if (chdir(argv[1]) != 0) {
        std::perror("chdir");
        return 2;
    }
    try {
        jsonChecks();
        settingsChecks(argc, argv);
        bcChecks();
        fieldIoChecks();
    } catch (const std::exception& e) {
        std::printf("FAIL  unexpected exception: %s\n", e.what());
        ++g_fail;
    }
    std::printf("%d of %d checks passed\n", g_total - g_fail, g_total);
    return g_fail ? 1 : 0;
}
