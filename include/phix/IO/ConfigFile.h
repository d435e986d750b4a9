#pragma once
// JSONC settings file; drop-in for include/IO/ConfigFile.h (PhiX::IO::ConfigFile)
// with the library's own JSON value (IO/Json.h) in place of nlohmann::json.
#include "IO/Json.h"

#include <string>

namespace PhiX {
namespace IO {

class ConfigFile {
public:
    // Loads and parses a JSONC file; throws std::runtime_error on failure.
    explicit ConfigFile(const std::string& path);

    // argv[1] if given, else defaultPath; prints the error and a usage line
    // and exits with status 1 on failure (ConfigFile.cpp fromArgs).
    static ConfigFile fromArgs(int argc, char* argv[], const std::string& defaultPath = "settings/settings.jsonc");

    // Top-level key (std::out_of_range if absent).
    const Json& operator[](const std::string& key) const { return data_.at(key); }
    bool has(const std::string& key) const { return data_.contains(key); }
    const Json& data() const { return data_; }

private:
    Json data_;
};

}  // namespace IO
}  // namespace PhiX
