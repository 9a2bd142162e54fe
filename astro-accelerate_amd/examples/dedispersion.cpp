// A C++ caller of the library through the reference's public headers, in the
// shape of the reference's examples/src/dedispersion.cpp:25-74: read a
// SIGPROC file, build an aa_ddtr_plan, run aa_pipeline_api<unsigned short>
// with copy_ddtr_data_to_host, and take the DM-time planes from
// output_buffer().
//
//   examples_dedispersion <file.fil> <out.bin> [low high step inBin]...
//
// Without ranges the plan is 0-370/0.307/1, 370-740/0.652/1,
// 740-1480/1.266/4, 1480-2950/25.12/8, 2950-5000/4.0/16.  out.bin holds,
// per range: int32 ndms, int32 samples, int32 inBin, float32 low, float32
// step, then ndms x samples float32 values ([dm][t]).
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <vector>

#include "aa_ddtr_plan.hpp"
#include "aa_ddtr_strategy.hpp"
#include "aa_device_info.hpp"
#include "aa_filterbank_metadata.hpp"
#include "aa_log.hpp"
#include "aa_pipeline_api.hpp"
#include "aa_sigproc_input.hpp"

using namespace astroaccelerate;

int main(int argc, char *argv[]) {
    if (argc < 3 || (argc - 3) % 4 != 0) {
        std::cerr << "usage: " << argv[0] << " <file.fil> <out.bin> [low high step inBin]...\n";
        return 2;
    }
    aa_ddtr_plan ddtr_plan;
    if (argc == 3) {
        ddtr_plan.add_dm(0, 370, 0.307, 1, 1);
        ddtr_plan.add_dm(370, 740, 0.652, 1, 1);
        ddtr_plan.add_dm(740, 1480, 1.266, 4, 4);
        ddtr_plan.add_dm(1480, 2950, 25.12, 8, 8);
        ddtr_plan.add_dm(2950, 5000, 4.000, 16, 16);
    } else {
        for (int a = 3; a < argc; a += 4) {
            const int bin = std::atoi(argv[a + 3]);
            ddtr_plan.add_dm((float)std::atof(argv[a]), (float)std::atof(argv[a + 1]), (float)std::atof(argv[a + 2]),
                             bin, bin);
        }
    }

    aa_sigproc_input filterbank_datafile(argv[1]);
    aa_filterbank_metadata metadata = filterbank_datafile.read_metadata();
    if (!filterbank_datafile.read_signal()) {
        LOG(log_level::error, "could not read the telescope data");
        return 1;
    }

    aa_device_info selected_device(0);
    aa_pipeline::pipeline pipeline_components;
    pipeline_components.insert(aa_pipeline::component::dedispersion);
    aa_pipeline::pipeline_option pipeline_options;
    pipeline_options.insert(aa_pipeline::component_option::copy_ddtr_data_to_host);

    aa_pipeline_api<unsigned short> pipeline_runner(pipeline_components, pipeline_options, metadata,
                                                    filterbank_datafile.input_buffer().data(), selected_device);
    if (!pipeline_runner.bind(ddtr_plan)) {
        LOG(log_level::error, "could not bind the plan");
        return 1;
    }
    if (!pipeline_runner.ready()) {
        LOG(log_level::error, "pipeline is not ready");
        return 1;
    }
    aa_pipeline_runner::status status_code;
    while (pipeline_runner.run(status_code)) {
    }
    if (status_code == aa_pipeline_runner::status::error) {
        LOG(log_level::error, "pipeline failed");
        return 1;
    }

    float ***ddtr_output = pipeline_runner.output_buffer();
    aa_ddtr_strategy plan = pipeline_runner.ddtr_strategy();
    const size_t nsamples = plan.nProcessedTimesamples();
    FILE *out = std::fopen(argv[2], "wb");
    if (!out) {
        LOG(log_level::error, "cannot write the output file");
        return 1;
    }
    bool ok = true;
    for (size_t r = 0; r < plan.get_nRanges(); ++r) {
        const aa_ddtr_plan::dm range = plan.dm(r);
        const int ndms = plan.ndms(r);
        const int samples = (int)(nsamples / (size_t)range.inBin);
        const int head[3] = {ndms, samples, range.inBin};
        const float dm[2] = {range.low, range.step};
        ok = ok && std::fwrite(head, sizeof(int), 3, out) == 3 && std::fwrite(dm, sizeof(float), 2, out) == 2;
        for (int d = 0; d < ndms && ok; ++d)
            ok = std::fwrite(ddtr_output[r][d], sizeof(float), (size_t)samples, out) == (size_t)samples;
        std::cout << "range " << r << ": DM " << range.low << " - " << range.high << " step " << range.step
                  << ", " << ndms << " trials x " << samples << " samples (bin " << range.inBin << ")\n";
    }
    ok = (std::fclose(out) == 0) && ok;
    pipeline_runner.cleanup();
    std::cout << (ok ? "NOTICE: Finished." : "ERROR: writing the output failed.") << std::endl;
    return ok ? 0 : 1;
}
