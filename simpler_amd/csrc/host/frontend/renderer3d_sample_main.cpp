// renderer3d_sample key=value ...: the reference's argv-driven renderer
// (reference renderer/samples/renderer3d_sample.cpp:41-870) on the gfx950
// engine.  Writes <outFilePrefix>.pfm3d (default USOCTRendering.pfm3d).
// With SIMPLER_DESCRIBE=<file> set, it writes the run's engine descriptor
// there as JSON instead of rendering (no GPU needed).
#include <cstdlib>
#include <iostream>

#include "settings.h"

int main(int argc, char **argv) {
    using namespace simpler_frontend;
    Settings s;
    if (settings_from_argv(argc, argv, s, std::cerr) != kOk) return kFail;
    if (s.print_inputs) {
        print_settings(std::cout, s);
        std::cout << "Total medium length = " << s.medium_r.x - s.medium_l.x << std::endl;
    }
    const char *describe = std::getenv("SIMPLER_DESCRIBE");
    return run(s, s.output_file_name, describe ? describe : "", std::cerr);
}
