// Facts of a model that decide which kernel shapes can run it (test infrastructure): for each
// variant line of a job (the emu_run format), one line of key=value pairs from the product's
// own model compiler and blob packer.
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>

#include "front.h"
#include "model.h"
#include "pack.h"

namespace {
std::string slurp(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot read " + path);
    std::stringstream ss;
    ss << f.rdbuf();
    return ss.str();
}
std::vector<float> floats(const std::string& path) {
    if (path == "-") return {};
    const std::string s = slurp(path);
    std::vector<float> v(s.size() / 4);
    std::memcpy(v.data(), s.data(), v.size() * 4);
    return v;
}
}  // namespace

int main(int argc, char** argv) {
    if (argc != 2) {
        std::cerr << "usage: emu_plan JOB\n";
        return 2;
    }
    try {
        std::ifstream job(argv[1]);
        std::string line;
        while (std::getline(job, line)) {
            std::istringstream ls(line);
            std::string key;
            if (!(ls >> key) || key != "variant") continue;
            std::string bnd, cfg, mu, md, is;
            uint64_t n, count, first, seed;
            uint32_t vid;
            ls >> bnd >> cfg >> n >> count >> first >> seed >> vid >> mu >> md >> is;
            mbg::ParsedModel pm;
            mbg::parse_model(slurp(bnd), slurp(cfg), pm);
            const mbg::CompiledModel cm = mbg::compile_model(pm);
            const std::vector<float> fu = floats(mu), fd = floats(md), fi = floats(is);
            const mbg::RateTables rt = mbg::make_rate_tables(cm, fu.empty() ? nullptr : fu.data(),
                                                             fd.empty() ? nullptr : fd.data(),
                                                             fi.empty() ? nullptr : fi.data());
            double deps = 0;
            for (const mbg::NodeRec& r : cm.nodes) deps += r.dep_cnt;
            const mbg::RunScales rs = mbg::run_scales(cm, n);
            std::cout << "N=" << cm.N << " NW=" << cm.NW << " K=" << cm.K << " W=" << rs.W
                      << " discrete=" << (cm.discrete ? 1 : 0) << " generic=" << (cm.any_generic ? 1 : 0)
                      << " prog=" << (cm.any_prog ? 1 : 0) << " max_in=" << cm.max_inputs
                      << " rate32=" << (mbg::rates_fit_u32(rt) ? 1 : 0) << " wsmad=" << (mbg::rates_fit_i31(rt) ? 1 : 0)
                      << " compact=" << (mbg::compact_fits(cm) ? 1 : 0)
                      << " blob=" << mbg::pack_blob(cm, rt, mbg::rates_fit_u32(rt), mbg::compact_fits(cm)).lay.bytes
                      << " mean_deps=" << (cm.N ? deps / cm.N : 0.0) << "\n";
        }
    } catch (const std::exception& e) {
        std::cerr << "emu_plan: " << e.what() << "\n";
        return 1;
    }
    return 0;
}
