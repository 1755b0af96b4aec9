// MaBoSSG command line: MaBoSSG [-o prefix] bnd_file cfg_file
//
// Same arguments, exit codes and outputs as the reference CLI (src/main.cpp:101-169): with -o it
// writes <prefix>_probtraj.csv and <prefix>_fp.csv, otherwise it prints the final states,
// fixed points and windows to stdout.  Exit status 1 on usage, parse, limit or device errors.
// Runs entirely through the C ABI of libmbg.so (include/mbg.h).
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>
#include <vector>

#include "../../include/mbg.h"

int main(int argc, char** argv) {
    std::vector<std::string> args(argv + 1, argv + argc);
    if (!(args.size() == 2 || (args.size() == 4 && args[0] == "-o"))) {
        std::cout << "Usage: MaBoSSG [-o prefix] bnd_file cfg_file" << std::endl;
        return 1;
    }
    std::string prefix, bnd = args[0], cfg = args[1];
    if (args[0] == "-o") {
        prefix = args[1];
        bnd = args[2];
        cfg = args[3];
    }
    char err[4096] = {0};
    mbg_model* model = nullptr;
    if (mbg_model_parse_files(bnd.c_str(), cfg.c_str(), &model, err, sizeof(err)) != MBG_OK) {
        std::cerr << err << std::endl;
        return 1;
    }
    double sample_count = 0;
    mbg_model_constant(model, "sample_count", &sample_count);
    mbg_ctx* ctx = nullptr;
    if (mbg_create(0, &ctx) != MBG_OK) {
        std::cerr << "no HIP device available" << std::endl;
        mbg_model_free(model);
        return 1;
    }
    mbg_run_params p = {};
    p.n_trajectories = (uint64_t)(int)sample_count;  // src/main.cpp:131 (float constant to int)
    p.use_model_seed = 1;
    // MBG_BATCH_RESUME=1 reproduces the reference's 100-record launches (a fresh waiting time
    // for the same state after every 99th step) instead of uninterrupted Gillespie
    const char* br = std::getenv("MBG_BATCH_RESUME");
    p.compat_batch_resume = br && std::atoi(br) != 0 ? 1 : 0;
    int rc = 0;
    if (mbg_run(ctx, model, &p) != MBG_OK) {
        std::cerr << mbg_last_error(ctx) << std::endl;
        rc = 1;
    } else if (!prefix.empty()) {
        if (mbg_write_csv(ctx, model, prefix.c_str()) != MBG_OK) {
            std::cerr << mbg_last_error(ctx) << std::endl;
            rc = 1;
        }
    } else if (mbg_write_text(ctx, model, "-") != MBG_OK) {
        std::cerr << mbg_last_error(ctx) << std::endl;
        rc = 1;
    }
    mbg_destroy(ctx);
    mbg_model_free(model);
    return rc;
}
