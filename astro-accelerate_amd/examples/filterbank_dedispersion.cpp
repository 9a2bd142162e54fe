// The C-style entry point dedisperse_telescope_data, called as the
// reference's examples/src/filterbank_dedispersion.cpp calls it (lines
// 26-48 there): read a SIGPROC file, list the DM ranges, run the
// dedispersion-only pipeline with the zero-DM option.
//
//   examples_filterbank_dedispersion <file.fil> [low high step inBin]...
//
// Without ranges the 8-range plan 0-150/0.1/1 ... 2000-3000/2.0/8 is used.
#include <cstdlib>
#include <iostream>
#include <vector>

#include "aa_pipeline_wrapper_functions.hpp"
#include "aa_sigproc_input.hpp"

using namespace astroaccelerate;

int main(int argc, char *argv[]) {
    if (argc < 2 || (argc - 2) % 4 != 0) {
        std::cerr << "usage: " << argv[0] << " <file.fil> [low high step inBin]...\n";
        return 2;
    }
    aa_sigproc_input filterbank_datafile(argv[1]);
    aa_filterbank_metadata filterbank_metadata = filterbank_datafile.read_metadata();
    if (!filterbank_datafile.read_signal()) {
        std::cout << "ERROR: Could not read telescope data." << std::endl;
        return 1;
    }
    std::vector<aa_ddtr_plan::dm> dm_ranges;
    if (argc == 2) {
        dm_ranges = {{0, 150, 0.1f, 1, 1},    {150, 300, 0.2f, 1, 1},   {300, 500, 0.25f, 1, 1},
                     {500, 900, 0.4f, 2, 2},  {900, 1200, 0.6f, 4, 4},  {1200, 1500, 0.8f, 4, 4},
                     {1500, 2000, 1.0f, 4, 4}, {2000, 3000, 2.0f, 8, 8}};
    } else {
        for (int a = 2; a < argc; a += 4) {
            const int bin = std::atoi(argv[a + 3]);
            dm_ranges.push_back({(float)std::atof(argv[a]), (float)std::atof(argv[a + 1]),
                                 (float)std::atof(argv[a + 2]), bin, bin});
        }
    }
    const aa_pipeline::pipeline_option pipeline_options = {aa_pipeline::component_option::zero_dm};
    dedisperse_telescope_data(filterbank_metadata, pipeline_options, dm_ranges, filterbank_datafile.input_buffer());
    std::cout << "NOTICE: Finished." << std::endl;
    return 0;
}
