// The command-line reconstruction (frontend/src/main.cpp + Framework): read a
// parameter file, open the projection stack, reconstruct on the GPU, write the
// volume.
#pragma once

#include <ostream>
#include <string>

#include "config.hpp"

namespace etn {

// Runs one reconstruction described by `c`; progress lines go to `log`.
void run_reconstruction(const ReconstructionConfig& c, std::ostream& log);

// main(argc, argv) of the reference's command-line tool, returning its exit
// code: 0 after a reconstruction, help text or a missing-parameter report
// without --config (main.cpp:98-107); 1 for an error; 2 for an unknown failure.
int framework_main(int argc, const char* const* argv, std::ostream& log);

// Parses argv (and the --config file it names) into the resolved parameter
// sets without running anything.
ReconstructionConfig resolve_config(int argc, const char* const* argv);

}  // namespace etn
