// config.hip -- the configuration and output layer of the drop-in C++ API
// (host code): JSON values and the JSONC parser, ConfigFile, the IO:: field
// file functions with restart resolution, OutputWriter, buildBCs.
// Reference interfaces: include/IO/ConfigFile.h, include/IO/FieldIO.h,
// include/IO/OutputWriter.h, include/boundary/BCFactory.h.
#include "IO/ConfigFile.h"
#include "IO/FieldIO.h"
#include "IO/OutputWriter.h"
#include "boundary/BCFactory.h"
#include "boundary/NoFluxBC.h"
#include "boundary/PeriodicBC.h"
#include "field/ScalarField.h"
#include "field/VectorField.h"

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <stdexcept>

namespace PhiX {
namespace IO {

// ------------------------------------------------------------------ JSON ----
class JsonParser {
public:
    explicit JsonParser(const std::string& text) : s_(text) {}

    Json document() {
        Json v = value();
        ws();
        if (p_ != s_.size()) fail("unexpected characters after the value");
        return v;
    }

private:
    const std::string& s_;
    std::size_t p_ = 0;
    int depth_ = 0;

    [[noreturn]] void fail(const std::string& what) const {
        int line = 1, col = 1;
        for (std::size_t i = 0; i < p_ && i < s_.size(); ++i) {
            if (s_[i] == '\n') {
                ++line;
                col = 1;
            } else {
                ++col;
            }
        }
        throw std::runtime_error("JSON parse error at line " + std::to_string(line) + ", column " +
                                 std::to_string(col) + ": " + what);
    }
    bool more() const { return p_ < s_.size(); }
    char peek() const { return more() ? s_[p_] : '\0'; }

    // whitespace and `//` line comments
    void ws() {
        while (more()) {
            const char c = s_[p_];
            if (c == ' ' || c == '\t' || c == '\r' || c == '\n') {
                ++p_;
            } else if (c == '/' && p_ + 1 < s_.size() && s_[p_ + 1] == '/') {
                while (more() && s_[p_] != '\n') ++p_;
            } else {
                break;
            }
        }
    }
    void expect(char c) {
        ws();
        if (peek() != c) fail(std::string("expected '") + c + "'");
        ++p_;
    }

    Json value() {
        ws();
        if (!more()) fail("unexpected end of input");
        if (++depth_ > 512) fail("nesting deeper than 512 levels");
        Json v;
        switch (peek()) {
            case '{': v = object(); break;
            case '[': v = array(); break;
            case '"':
                v.type_ = Json::Type::String;
                v.str_ = str();
                break;
            case 't': word("true"); v.type_ = Json::Type::Boolean; v.bool_ = true; break;
            case 'f': word("false"); v.type_ = Json::Type::Boolean; v.bool_ = false; break;
            case 'n': word("null"); break;
            default: v = number();
        }
        --depth_;
        return v;
    }

    void word(const char* w) {
        const std::size_t n = std::char_traits<char>::length(w);
        if (s_.compare(p_, n, w) != 0) fail(std::string("invalid literal (expected ") + w + ")");
        p_ += n;
    }

    Json object() {
        ++p_;   // '{'
        Json o;
        o.type_ = Json::Type::Object;
        ws();
        if (peek() == '}') {
            ++p_;
            return o;
        }
        for (;;) {
            ws();
            if (peek() != '"') fail("expected a string key");
            std::string key = str();
            expect(':');
            Json v = value();
            auto it = std::find_if(o.obj_.begin(), o.obj_.end(), [&](const auto& kv) { return kv.first == key; });
            if (it != o.obj_.end()) it->second = std::move(v);
            else o.obj_.emplace_back(std::move(key), std::move(v));
            ws();
            if (peek() == ',') {
                ++p_;
                continue;
            }
            if (peek() == '}') {
                ++p_;
                return o;
            }
            fail("expected ',' or '}' in an object");
        }
    }

    Json array() {
        ++p_;   // '['
        Json a;
        a.type_ = Json::Type::Array;
        ws();
        if (peek() == ']') {
            ++p_;
            return a;
        }
        for (;;) {
            a.arr_.push_back(value());
            ws();
            if (peek() == ',') {
                ++p_;
                continue;
            }
            if (peek() == ']') {
                ++p_;
                return a;
            }
            fail("expected ',' or ']' in an array");
        }
    }

    static void utf8(std::string& out, unsigned cp) {
        if (cp < 0x80) {
            out += (char)cp;
        } else if (cp < 0x800) {
            out += (char)(0xC0 | (cp >> 6));
            out += (char)(0x80 | (cp & 0x3F));
        } else if (cp < 0x10000) {
            out += (char)(0xE0 | (cp >> 12));
            out += (char)(0x80 | ((cp >> 6) & 0x3F));
            out += (char)(0x80 | (cp & 0x3F));
        } else {
            out += (char)(0xF0 | (cp >> 18));
            out += (char)(0x80 | ((cp >> 12) & 0x3F));
            out += (char)(0x80 | ((cp >> 6) & 0x3F));
            out += (char)(0x80 | (cp & 0x3F));
        }
    }
    unsigned hex4() {
        if (p_ + 4 > s_.size()) fail("truncated \\u escape");
        unsigned v = 0;
        for (int i = 0; i < 4; ++i) {
            const char c = s_[p_++];
            v <<= 4;
            if (c >= '0' && c <= '9') v |= (unsigned)(c - '0');
            else if (c >= 'a' && c <= 'f') v |= (unsigned)(c - 'a' + 10);
            else if (c >= 'A' && c <= 'F') v |= (unsigned)(c - 'A' + 10);
            else fail("invalid hex digit in \\u escape");
        }
        return v;
    }

    std::string str() {
        ++p_;   // opening quote
        std::string out;
        for (;;) {
            if (!more()) fail("unterminated string");
            const char c = s_[p_++];
            if (c == '"') return out;
            if ((unsigned char)c < 0x20) fail("control character in a string");
            if (c != '\\') {
                out += c;
                continue;
            }
            if (!more()) fail("unterminated escape");
            const char e = s_[p_++];
            switch (e) {
                case '"': out += '"'; break;
                case '\\': out += '\\'; break;
                case '/': out += '/'; break;
                case 'b': out += '\b'; break;
                case 'f': out += '\f'; break;
                case 'n': out += '\n'; break;
                case 'r': out += '\r'; break;
                case 't': out += '\t'; break;
                case 'u': {
                    unsigned cp = hex4();
                    if (cp >= 0xD800 && cp < 0xDC00) {   // high surrogate: a low one must follow
                        if (s_.compare(p_, 2, "\\u") != 0) fail("unpaired surrogate in \\u escape");
                        p_ += 2;
                        const unsigned lo = hex4();
                        if (lo < 0xDC00 || lo >= 0xE000) fail("invalid low surrogate in \\u escape");
                        cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                    } else if (cp >= 0xDC00 && cp < 0xE000) {
                        fail("unpaired surrogate in \\u escape");
                    }
                    utf8(out, cp);
                    break;
                }
                default: fail(std::string("invalid escape \\") + e);
            }
        }
    }

    // JSON number grammar: -?(0|[1-9][0-9]*)(\.[0-9]+)?([eE][+-]?[0-9]+)?
    Json number() {
        const std::size_t b = p_;
        auto digits = [&] {
            const std::size_t d = p_;
            while (more() && s_[p_] >= '0' && s_[p_] <= '9') ++p_;
            return p_ > d;
        };
        if (peek() == '-') ++p_;
        if (peek() == '0') ++p_;
        else if (!digits()) fail("invalid value");
        bool integral = true;
        if (peek() == '.') {
            ++p_;
            integral = false;
            if (!digits()) fail("digits expected after '.'");
        }
        if (peek() == 'e' || peek() == 'E') {
            ++p_;
            integral = false;
            if (peek() == '+' || peek() == '-') ++p_;
            if (!digits()) fail("digits expected in the exponent");
        }
        const std::string tok = s_.substr(b, p_ - b);
        Json n;
        n.type_ = Json::Type::Number;
        n.num_ = std::strtod(tok.c_str(), nullptr);
        if (integral) {
            errno = 0;
            const long long v = std::strtoll(tok.c_str(), nullptr, 10);
            if (errno != ERANGE) {
                n.integral_ = true;
                n.int_ = v;
            }
        }
        return n;
    }
};

Json Json::parse(const std::string& text) { return JsonParser(text).document(); }

void Json::typeError(const char* want) const {
    static const char* names[] = {"null", "boolean", "number", "string", "array", "object"};
    throw std::runtime_error(std::string("JSON: value is ") + names[(int)type_] + ", expected " + want);
}

const Json& Json::at(const std::string& key) const {
    if (type_ != Type::Object) typeError("an object");
    for (const auto& kv : obj_)
        if (kv.first == key) return kv.second;
    throw std::out_of_range("JSON: key \"" + key + "\" not found");
}

bool Json::contains(const std::string& key) const {
    if (type_ != Type::Object) return false;
    for (const auto& kv : obj_)
        if (kv.first == key) return true;
    return false;
}

const std::vector<std::pair<std::string, Json>>& Json::items() const {
    if (type_ != Type::Object) typeError("an object");
    return obj_;
}

const Json& Json::at(std::size_t i) const {
    if (type_ != Type::Array) typeError("an array");
    if (i >= arr_.size()) throw std::out_of_range("JSON: index " + std::to_string(i) + " out of range");
    return arr_[i];
}

std::size_t Json::size() const {
    if (type_ == Type::Array) return arr_.size();
    if (type_ == Type::Object) return obj_.size();
    return type_ == Type::Null ? 0 : 1;
}

double Json::asDouble() const {
    if (type_ != Type::Number) typeError("a number");
    return integral_ ? (double)int_ : num_;
}
// non-integral numbers truncate toward zero, as nlohmann::json converts them
long long Json::asInteger() const {
    if (type_ != Type::Number) typeError("a number");
    return integral_ ? int_ : (long long)num_;
}
const std::string& Json::asString() const {
    if (type_ != Type::String) typeError("a string");
    return str_;
}
bool Json::asBool() const {
    if (type_ != Type::Boolean) typeError("a boolean");
    return bool_;
}

static void dumpString(std::ostringstream& os, const std::string& s) {
    os << '"';
    for (const char c : s) {
        switch (c) {
            case '"': os << "\\\""; break;
            case '\\': os << "\\\\"; break;
            case '\n': os << "\\n"; break;
            case '\r': os << "\\r"; break;
            case '\t': os << "\\t"; break;
            case '\b': os << "\\b"; break;
            case '\f': os << "\\f"; break;
            default:
                if ((unsigned char)c < 0x20) {
                    char buf[8];
                    std::snprintf(buf, sizeof buf, "\\u%04x", (unsigned)c);
                    os << buf;
                } else {
                    os << c;
                }
        }
    }
    os << '"';
}

static void dumpValue(std::ostringstream& os, const Json& j) {
    switch (j.type()) {
        case Json::Type::Null: os << "null"; break;
        case Json::Type::Boolean: os << (j.get<bool>() ? "true" : "false"); break;
        case Json::Type::Number:
            if (j.is_number_integer()) os << j.get<long long>();
            else os << std::setprecision(17) << j.get<double>();
            break;
        case Json::Type::String: dumpString(os, j.get<std::string>()); break;
        case Json::Type::Array:
            os << '[';
            for (std::size_t i = 0; i < j.size(); ++i) {
                if (i) os << ',';
                dumpValue(os, j.at(i));
            }
            os << ']';
            break;
        case Json::Type::Object: {
            os << '{';
            bool first = true;
            for (const auto& kv : j.items()) {
                if (!first) os << ',';
                first = false;
                dumpString(os, kv.first);
                os << ':';
                dumpValue(os, kv.second);
            }
            os << '}';
        }
    }
}

std::string Json::dump() const {
    std::ostringstream os;
    dumpValue(os, *this);
    return os.str();
}

// ------------------------------------------------------------ ConfigFile ----
ConfigFile::ConfigFile(const std::string& path) {
    std::ifstream is(path);
    if (!is.is_open()) throw std::runtime_error("ConfigFile: cannot open \"" + path + "\"");
    std::ostringstream text;
    text << is.rdbuf();
    try {
        data_ = Json::parse(text.str());
    } catch (const std::runtime_error& e) {
        throw std::runtime_error("ConfigFile: " + std::string(e.what()) + " in \"" + path + "\"");
    }
}

ConfigFile ConfigFile::fromArgs(int argc, char* argv[], const std::string& defaultPath) {
    const std::string path = argc >= 2 ? argv[1] : defaultPath;
    try {
        return ConfigFile(path);
    } catch (const std::exception& e) {
        std::cerr << e.what() << "\n  Usage: " << (argc >= 1 ? argv[0] : "solver") << " [path/to/settings.jsonc]\n";
        std::exit(1);
    }
}

// --------------------------------------------------------------- FieldIO ----
namespace {

// First line of a field file: which kind and format it is.
std::string firstLine(const std::string& path, const char* who) {
    std::ifstream is(path, std::ios::binary);
    if (!is) throw std::runtime_error(std::string(who) + ": cannot open file: " + path);
    std::string line;
    std::getline(is, line);
    if (!line.empty() && line.back() == '\r') line.pop_back();
    return line;
}

bool isDat(const std::string& first) { return first.find("- DAT") != std::string::npos; }

// DAT files (FieldIO.cpp:114-192): '#' header lines carrying the tokens
// "name: <name>", "nComponents: N" (vector fields) and "nx N ny N nz N" in any
// order, then one line per physical cell, x fastest: x y z v0 [v1 ...].
struct DatFile {
    std::string name;
    int nx = 0, ny = 0, nz = 0, ncomp = 0;
    std::vector<std::vector<double>> comp;
};

DatFile readDat(const std::string& path, bool vector) {
    std::ifstream is(path);
    if (!is) throw std::runtime_error("IO::readField: cannot open file: " + path);
    DatFile d;
    std::string line;
    std::streampos data_at = is.tellg();
    while (std::getline(is, line)) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        if (line.empty() || line[0] != '#') break;
        data_at = is.tellg();
        std::istringstream ss(line.substr(1));
        std::string tok;
        while (ss >> tok) {
            if (tok == "name:") ss >> d.name;
            else if (tok == "nComponents:") ss >> d.ncomp;
            else if (tok == "nx") ss >> d.nx;
            else if (tok == "ny") ss >> d.ny;
            else if (tok == "nz") ss >> d.nz;
        }
    }
    if (!vector) d.ncomp = 1;
    if (d.ncomp <= 0) throw std::runtime_error("IO::readVectorField: invalid nComponents in " + path);
    if (d.nx <= 0 || d.ny <= 0 || d.nz <= 0) throw std::runtime_error("IO::readField: missing nx/ny/nz in " + path);
    is.clear();
    is.seekg(data_at);
    const std::size_t n = (std::size_t)d.nx * d.ny * d.nz;
    d.comp.assign(d.ncomp, std::vector<double>(n));
    for (std::size_t q = 0; q < n; ++q) {
        if (!std::getline(is, line)) throw std::runtime_error("IO::readField: unexpected end of DAT data in " + path);
        std::istringstream ss(line);
        double x, y, z;
        if (!(ss >> x >> y >> z)) throw std::runtime_error("IO::readField: malformed DAT line in " + path);
        for (int c = 0; c < d.ncomp; ++c)
            if (!(ss >> d.comp[c][q])) throw std::runtime_error("IO::readField: malformed DAT line in " + path);
    }
    return d;
}

void checkMesh(int nx, int ny, int nz, const Mesh& m, const std::string& what, const std::string& path) {
    if (nx != m.n[0] || ny != m.n[1] || nz != m.n[2])
        throw std::runtime_error("IO::readField: file dimensions (" + std::to_string(nx) + "x" + std::to_string(ny) +
                                 "x" + std::to_string(nz) + ") do not match " + what + " mesh (" +
                                 std::to_string(m.n[0]) + "x" + std::to_string(m.n[1]) + "x" +
                                 std::to_string(m.n[2]) + ") in " + path);
}

// Copies the physical cells of src into dst (same mesh).
void copyPhysical(ScalarField& dst, const ScalarField& src) {
    for (int k = 0; k < dst.mesh.n[2]; ++k)
        for (int j = 0; j < dst.mesh.n[1]; ++j)
            for (int i = 0; i < dst.mesh.n[0]; ++i) dst.curr[dst.index(i, j, k)] = src.curr[src.index(i, j, k)];
}

void fillPhysical(ScalarField& dst, const std::vector<double>& v) {
    std::size_t q = 0;
    for (int k = 0; k < dst.mesh.n[2]; ++k)
        for (int j = 0; j < dst.mesh.n[1]; ++j)
            for (int i = 0; i < dst.mesh.n[0]; ++i) dst.curr[dst.index(i, j, k)] = v[q++];
}

}  // namespace

void writeField(const ScalarField& f, const std::string& path, FieldFormat fmt) { f.write(path, fmt); }

ScalarField readScalarField(const Mesh& mesh, const std::string& path, int ghost) {
    if (!isDat(firstLine(path, "IO::readScalarField"))) return ScalarField::readFromFile(mesh, path, ghost);
    const DatFile d = readDat(path, false);
    checkMesh(d.nx, d.ny, d.nz, mesh, "the provided", path);
    ScalarField f(mesh, d.name, ghost);
    fillPhysical(f, d.comp[0]);
    return f;
}

void readField(ScalarField& f, const std::string& path) {
    if (f.curr.size() != f.storedSize) throw std::runtime_error("IO::readField: field has no host storage");
    if (isDat(firstLine(path, "IO::readField"))) {
        const DatFile d = readDat(path, false);
        checkMesh(d.nx, d.ny, d.nz, f.mesh, "field \"" + f.name + "\"", path);
        fillPhysical(f, d.comp[0]);
        return;
    }
    const ScalarField tmp = ScalarField::readFromFile(f.mesh, path, f.ghost);
    copyPhysical(f, tmp);
}

void writeField(const VectorField& vf, const std::string& path, FieldFormat fmt) { vf.write(path, fmt); }

VectorField readVectorField(const Mesh& mesh, const std::string& path, int ghost) {
    if (!isDat(firstLine(path, "IO::readVectorField"))) return VectorField::readFromFile(mesh, path, ghost);
    const DatFile d = readDat(path, true);
    checkMesh(d.nx, d.ny, d.nz, mesh, "the provided", path);
    VectorField vf(mesh, d.name, d.ncomp, ghost);
    for (int c = 0; c < d.ncomp; ++c) fillPhysical(vf[c], d.comp[c]);
    return vf;
}

void readField(VectorField& vf, const std::string& path) {
    if (isDat(firstLine(path, "IO::readField"))) {
        const DatFile d = readDat(path, true);
        checkMesh(d.nx, d.ny, d.nz, vf.mesh, "field \"" + vf.name + "\"", path);
        if (d.ncomp != vf.nComponents())
            throw std::runtime_error("IO::readField: component count in " + path + " does not match the field");
        for (int c = 0; c < d.ncomp; ++c) fillPhysical(vf[c], d.comp[c]);
        return;
    }
    const VectorField tmp = VectorField::readFromFile(vf.mesh, path, vf.ghost);
    if (tmp.nComponents() != vf.nComponents())
        throw std::runtime_error("IO::readField: component count in " + path + " does not match the field");
    for (int c = 0; c < vf.nComponents(); ++c) copyPhysical(vf[c], tmp[c]);
}

int resolveStartStep(const std::string& startFrom, const std::string& refFieldName) {
    if (startFrom == "initial_field") return 0;
    // a leading step number; std::invalid_argument when there is none
    const int target = std::stoi(startFrom);
    namespace fs = std::filesystem;
    const std::string prefix = refFieldName + "_";
    std::vector<int> steps;
    if (fs::is_directory("output")) {
        for (const auto& e : fs::directory_iterator("output")) {
            if (e.path().extension() != ".field") continue;
            const std::string stem = e.path().stem().string();
            if (stem.compare(0, prefix.size(), prefix) != 0) continue;
            try {   // stems whose remainder does not start with a number are skipped
                steps.push_back(std::stoi(stem.substr(prefix.size())));
            } catch (const std::exception&) {
            }
        }
    }
    if (steps.empty())
        throw std::runtime_error("IO::resolveStartStep: start_from=\"" + startFrom + "\" but no " + prefix +
                                 "*.field files found in output/");
    // nearest step; ties go to the earlier one
    std::sort(steps.begin(), steps.end());
    int best = steps[0];
    for (const int s : steps)
        if (std::abs(s - target) < std::abs(best - target)) best = s;
    return best;
}

void initField(ScalarField& f, int startStep) {
    const std::string path = startStep == 0 ? "settings/initial_field/" + f.name + ".field"
                                            : "output/" + f.name + "_" + std::to_string(startStep) + ".field";
    std::cout << (startStep == 0 ? "  cold start: loading " : "  warm start: loading ") << path << "\n";
    readField(f, path);
}

// ---------------------------------------------------------- OutputWriter ----
OutputWriter::OutputWriter(const Json& cfg)
    : printInterval(cfg.at("print_interval").get<int>()), writeInterval(cfg.at("write_interval").get<int>()) {
    if (printInterval < 1 || writeInterval < 1)
        throw std::runtime_error("OutputWriter: print_interval and write_interval must be >= 1");
    const std::string fmt = cfg.at("format").get<std::string>();
    writeBinary_ = fmt == "BINARY" || fmt == "ALL";
    writeDat_ = fmt == "DAT" || fmt == "ALL";
    writeVtk_ = fmt == "VTK" || fmt == "ALL";   // any other name writes no files, as in the reference
    std::filesystem::create_directories("output");
    resetTimer();
}

void OutputWriter::writeFields(ScalarField& f, int step, double simTime) {
    if (f.deviceAllocated()) f.downloadCurrFromDevice();
    const std::string base = "output/" + f.name + "_" + std::to_string(step);
    if (writeBinary_) writeField(f, base + ".field", FieldFormat::BINARY);
    if (writeDat_) writeField(f, base + ".dat", FieldFormat::DAT);
    if (writeVtk_) writeField(f, base + ".vts", FieldFormat::VTS);
    std::cout << "  step " << step << "  t=" << simTime << "  written: " << base << "\n" << std::flush;
}

void OutputWriter::printProgress(int step, double simTime) {
    const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start_).count();
    std::ostringstream el;   // formats the seconds without changing std::cout's flags
    el << std::fixed << std::setprecision(1) << s;
    std::cout << "  [progress] step=" << step << "  t=" << simTime << "  elapsed=" << el.str() << "s\n" << std::flush;
}

void OutputWriter::resetTimer() { t_start_ = std::chrono::steady_clock::now(); }

}  // namespace IO

// ------------------------------------------------------------- buildBCs ----
namespace {

void addAxis(BCSet& set, Axis axis, const std::string& lo, const std::string& hi) {
    const bool plo = lo == "Periodic", phi = hi == "Periodic";
    if (plo != phi) {
        const char* name = axis == Axis::X ? "X" : (axis == Axis::Y ? "Y" : "Z");
        throw std::runtime_error(std::string("buildBCs: Periodic BC must be set on both sides of axis ") + name +
                                 ", got \"" + lo + "\" / \"" + hi + "\"");
    }
    if (plo) {
        set.storage.push_back(std::make_unique<PeriodicBC>(axis));
        set.ptrs.push_back(set.storage.back().get());
        return;
    }
    for (const auto& side : {std::make_pair(lo, Side::LOW), std::make_pair(hi, Side::HIGH)}) {
        if (side.first != "NoFlux") throw std::runtime_error("buildBCs: unsupported BC type \"" + side.first + "\"");
        set.storage.push_back(std::make_unique<NoFluxBC>(axis, side.second));
        set.ptrs.push_back(set.storage.back().get());
    }
}

}  // namespace

BCSet buildBCs(const IO::Json& cfg) {
    BCSet set;
    addAxis(set, Axis::X, cfg.at("x_min").get<std::string>(), cfg.at("x_max").get<std::string>());
    if (cfg.contains("y_min") && cfg.contains("y_max"))
        addAxis(set, Axis::Y, cfg.at("y_min").get<std::string>(), cfg.at("y_max").get<std::string>());
    if (cfg.contains("z_min") && cfg.contains("z_max"))
        addAxis(set, Axis::Z, cfg.at("z_min").get<std::string>(), cfg.at("z_max").get<std::string>());
    return set;
}

}  // namespace PhiX
