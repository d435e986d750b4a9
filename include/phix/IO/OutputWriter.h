#pragma once
// Output cadence and field writing from a settings block; drop-in for
// include/IO/OutputWriter.h:
//   { "print_interval": 10000, "write_interval": 100000,
//     "format": "BINARY" | "DAT" | "VTK" | "ALL" }
// Files go to output/<name>_<step>.field / .dat / .vts.
#include "IO/FieldFormat.h"
#include "IO/Json.h"
#include "field/ScalarField.h"

#include <chrono>
#include <string>

namespace PhiX {
namespace IO {

class OutputWriter {
public:
    explicit OutputWriter(const Json& output_config);

    int printInterval;
    int writeInterval;

    bool shouldPrint(int step) const { return step % printInterval == 0; }
    bool shouldWrite(int step) const { return step % writeInterval == 0; }

    // Downloads curr from the device (when allocated), writes every configured
    // format and prints a status line.
    void writeFields(ScalarField& f, int step, double simTime);
    // Step, simulation time and wall-clock seconds since the last reset.
    void printProgress(int step, double simTime);
    void resetTimer();

private:
    bool writeBinary_, writeDat_, writeVtk_;
    std::chrono::steady_clock::time_point t_start_;
};

}  // namespace IO
}  // namespace PhiX
