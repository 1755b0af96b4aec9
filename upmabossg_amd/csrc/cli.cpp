// MaBoSSG command line: MaBoSSG [-o prefix] bnd_file cfg_file
//
// Same arguments, exit codes and outputs as the reference CLI (src/main.cpp:101-169): with -o it
// writes <prefix>_probtraj.csv and <prefix>_fp.csv, otherwise it prints the final states,
// fixed points and windows to stdout.  Exit status 1 on usage, parse, limit or device errors.
// Runs entirely through the C ABI of libmbg.so (include/mbg.h).
//
// Extension: MaBoSSG [-o prefix] -v variants_file bnd_file cfg_file runs every variant of the
// file (format in include/mbg.h) as one batched sweep and writes, per variant NAME,
// <prefix>_NAME_probtraj.csv and <prefix>_NAME_fp.csv, or without -o a "variant NAME" line
// followed by the stdout format.  Each variant's output is the output of a plain run on the
// model with that variant's overrides.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>
#include <vector>

#include "../../include/mbg.h"

namespace {

int usage() {
    std::cout << "Usage: MaBoSSG [-o prefix] bnd_file cfg_file" << std::endl;
    std::cerr << "       MaBoSSG [-o prefix] -v variants_file bnd_file cfg_file" << std::endl;
    return 1;
}

// Window-table bytes of one run of model m (W x 2^K 64-bit cells): sweeps are split so that
// the tables of one launch stay within a fixed device budget.
double table_bytes(const mbg_model* m) {
    double max_time = 0, tick = 1;
    mbg_model_constant(m, "max_time", &max_time);
    mbg_model_constant(m, "time_tick", &tick);
    const double W = tick > 0 ? std::ceil(max_time / tick) + 1 : 1;
    return W * std::ldexp(8.0, mbg_model_num_readouts(m));
}

// Parses the variants file against the model: the variants (each with its own model, freed by
// the caller) and their names.
bool load_variants(const mbg_model* base, const std::string& path, std::vector<mbg_variant>& vs,
                   std::vector<std::string>& names) {
    char err[4096] = {0};
    int count = 0;
    if (mbg_variants_parse_file(base, path.c_str(), nullptr, nullptr, 0, &count, err, sizeof(err)) != MBG_OK) {
        std::cerr << err << std::endl;
        return false;
    }
    vs.resize(count);
    std::vector<char> buf((size_t)count * 64 + 1, 0);
    if (count > 0 &&
        mbg_variants_parse_file(base, path.c_str(), vs.data(), buf.data(), count, &count, err, sizeof(err)) != MBG_OK) {
        std::cerr << err << std::endl;
        vs.clear();
        return false;
    }
    for (int i = 0; i < count; i++) names.emplace_back(&buf[(size_t)i * 64]);
    return true;
}

// ErrorTH, HD=0 and ErrorProba are filled in (MaBoSS's standard errors of the window averages);
// MBG_ERROR_COLUMNS=0 writes zeros there, as the reference does.
int error_columns() {
    const char* ec = std::getenv("MBG_ERROR_COLUMNS");
    return ec && std::atoi(ec) == 0 ? 0 : 1;
}

// More than 20 readouts: sweeps keep dense tables only, so each variant is a plain run (same
// trajectory count, seed and draws as in a sweep).
int run_variants_one_by_one(mbg_ctx* ctx, const std::vector<mbg_variant>& vs, const std::vector<std::string>& names,
                            const std::string& prefix) {
    for (size_t i = 0; i < vs.size(); i++) {
        mbg_run_params p = {};
        p.n_trajectories = vs[i].n_trajectories;
        p.seed = vs[i].seed;
        p.use_model_seed = vs[i].use_model_seed;
        p.variant_id = vs[i].variant_id;
        p.error_columns = error_columns();
        if (mbg_run(ctx, vs[i].model, &p) != MBG_OK) {
            std::cerr << mbg_last_error(ctx) << std::endl;
            return 1;
        }
        const mbg_model* m = vs[i].model;
        mbg_status st;
        if (!prefix.empty()) {
            st = mbg_write_csv(ctx, m, (prefix + "_" + names[i]).c_str());
        } else {
            std::cout << "variant " << names[i] << std::endl;
            st = mbg_write_text(ctx, m, "-");
        }
        if (st != MBG_OK) {
            std::cerr << mbg_last_error(ctx) << std::endl;
            return 1;
        }
    }
    return 0;
}

int run_variants(mbg_ctx* ctx, const mbg_model* base, const std::vector<mbg_variant>& vs,
                 const std::vector<std::string>& names, const std::string& prefix) {
    if (mbg_model_num_readouts(base) > 20) return run_variants_one_by_one(ctx, vs, names, prefix);
    const int count = (int)vs.size();
    const double budget = 32.0 * (1ull << 30);
    const int group = std::max(1, std::min(256, (int)(budget / table_bytes(base))));
    for (int g0 = 0; g0 < count; g0 += group) {
        const int g = std::min(group, count - g0);
        if (mbg_run_variants_ex(ctx, vs.data() + g0, g, error_columns() ? MBG_SWEEP_ERROR_COLUMNS : 0) != MBG_OK) {
            std::cerr << mbg_last_error(ctx) << std::endl;
            return 1;
        }
        for (int i = 0; i < g; i++) {
            const std::string& name = names[(size_t)(g0 + i)];
            mbg_status st;
            if (!prefix.empty()) {
                st = mbg_write_variant_csv(ctx, i, (prefix + "_" + name).c_str());
            } else {
                std::cout << "variant " << name << std::endl;
                st = mbg_write_variant_text(ctx, i, "-");
            }
            if (st != MBG_OK) {
                std::cerr << mbg_last_error(ctx) << std::endl;
                return 1;
            }
        }
    }
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    std::vector<std::string> args(argv + 1, argv + argc);
    std::string prefix, variants, bnd, cfg;
    if (args.size() == 2) {
        bnd = args[0];
        cfg = args[1];
    } else if (args.size() == 4 && args[0] == "-o") {
        prefix = args[1];
        bnd = args[2];
        cfg = args[3];
    } else if (args.size() == 4 && args[0] == "-v") {
        variants = args[1];
        bnd = args[2];
        cfg = args[3];
    } else if (args.size() == 6 && ((args[0] == "-o" && args[2] == "-v") || (args[0] == "-v" && args[2] == "-o"))) {
        (args[0] == "-o" ? prefix : variants) = args[1];
        (args[2] == "-o" ? prefix : variants) = args[3];
        bnd = args[4];
        cfg = args[5];
    } else {
        return usage();
    }
    char err[4096] = {0};
    mbg_model* model = nullptr;
    if (mbg_model_parse_files(bnd.c_str(), cfg.c_str(), &model, err, sizeof(err)) != MBG_OK) {
        std::cerr << err << std::endl;
        return 1;
    }
    double sample_count = 0;
    mbg_model_constant(model, "sample_count", &sample_count);
    std::vector<mbg_variant> vs;
    std::vector<std::string> names;
    auto free_all = [&] {
        for (mbg_variant& v : vs) mbg_model_free(const_cast<mbg_model*>(v.model));
        mbg_model_free(model);
    };
    if (!variants.empty() && !load_variants(model, variants, vs, names)) {
        free_all();
        return 1;
    }
    mbg_ctx* ctx = nullptr;
    if (mbg_create(0, &ctx) != MBG_OK) {
        std::cerr << "no HIP device available" << std::endl;
        free_all();
        return 1;
    }
    int rc = 0;
    if (!variants.empty()) {
        rc = run_variants(ctx, model, vs, names, prefix);
    } else {
        mbg_run_params p = {};
        p.n_trajectories = (uint64_t)(int)sample_count;  // src/main.cpp:131 (float constant to int)
        p.use_model_seed = 1;
        // MBG_BATCH_RESUME=1 reproduces the reference's 100-record launches (a fresh waiting
        // time for the same state after every 99th step) instead of uninterrupted Gillespie
        const char* br = std::getenv("MBG_BATCH_RESUME");
        p.compat_batch_resume = br && std::atoi(br) != 0 ? 1 : 0;
        p.error_columns = error_columns();
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
    }
    mbg_destroy(ctx);
    free_all();
    return rc;
}
