// simplER <settings.json>: the reference's JSON-driven renderer
// (reference simplER.cpp:21-164) on the gfx950 engine.  Writes
// <rendering_type>_YYYY_MM_DD_HH_MM.pfm3d in the working directory.
// With SIMPLER_DESCRIBE=<file> set, it writes the run's engine descriptor
// there as JSON instead of rendering (no GPU needed).
#include <cstdlib>
#include <iostream>

#include "settings.h"

int main(int argc, char **argv) {
    using namespace simpler_frontend;
    const std::string config = (argc > 1) ? argv[1] : "config.json";
    std::cout << "Using config file: " << config << std::endl;
    Settings s;
    if (settings_from_json_file(config, s, std::cerr) != kOk) return kFail;
    if (s.print_inputs) {
        std::cout << "simplER renderer started with the following parameters:\n";
        print_settings(std::cout, s);
        std::cout.flush();
    }
    const char *describe = std::getenv("SIMPLER_DESCRIBE");
    return run(s, s.output_file_name, describe ? describe : "", std::cerr);
}
