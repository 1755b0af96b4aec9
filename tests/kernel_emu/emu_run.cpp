// CPU emulation of one run of the trajectory kernel, as test infrastructure: the kernel source
// generated for one launch shape (EMU_KERNEL, the defines plus mbg_math.h and mbg_kernel.hip) is
// compiled for the host against emu_device.h and driven the way the runtime drives it on the
// GPU (csrc/runtime.cpp run_impl, and run_variants_impl for MBG_MULTI): the initial-state pass
// into the staging rows, then the persistent simulation kernel over a chosen grid, then the raw
// statistics tables.  The model tables come from the product's own front end, model compiler
// and blob packer (front.cpp, model.cpp, pack.cpp).
//
// Usage: emu_run JOB OUTDIR
//   JOB lines:  grid G | compat 0/1 | cell_cap C | stride_extra E | err_cap M |
//               variant BND CFG N COUNT FIRST_ID SEED VARIANT_ID MU MD IS
//   (MU / MD / IS: files of N float32 values, or "-"; one variant line for a single run)
// Writes OUTDIR/v<i>_<name>.u64 (prob, th, fin, info, fp, summ, cells, cellv) per variant.
#include "emu_device.h"

// the generated kernel source for the launch shape under test
#include EMU_KERNEL

#include <fstream>
#include <iostream>
#include <memory>
#include <sstream>

#include "front.h"
#include "model.h"
#include "pack.h"

#if MBG_MULTI
static_assert(MBG_NW <= 128, "the emulated sweep uses mbg_init_rates_multi (at most 128 words)");
#endif

namespace {

std::string slurp(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot read " + path);
    std::stringstream ss;
    ss << f.rdbuf();
    std::string s = ss.str();
    emu_unpoison(s.data(), s.size());
    return s;
}

std::vector<float> floats(const std::string& path) {
    if (path == "-") return {};
    const std::string s = slurp(path);
    std::vector<float> v(s.size() / 4);
    std::memcpy(v.data(), s.data(), v.size() * 4);
    return v;
}

void dump(const std::string& path, const void* p, size_t bytes) {
    std::ofstream f(path, std::ios::binary);
    f.write(static_cast<const char*>(p), (std::streamsize)bytes);
    if (!f) throw std::runtime_error("cannot write " + path);
}

struct Variant {
    mbg::ParsedModel pm;
    mbg::CompiledModel cm;
    mbg::RateTables rt;
    mbg::Blob blob;
    mbg::RunScales rs;
    uint64_t n = 0, count = 0, first_id = 0, seed = 0;
    uint32_t vid = 0;
    // emulated device buffers, each exactly sized (bounds-checked under AddressSanitizer)
    std::vector<u64> stage, prob, th, fin, fp, summ, counters, cells, cellv, cell_miss, prob2, th2, cellv2;
    std::vector<float> istate;
    mbg_desc D{};
};

}  // namespace

int main(int argc, char** argv) {
    if (argc != 3) {
        std::cerr << "usage: emu_run JOB OUTDIR\n";
        return 2;
    }
    try {
        std::ifstream job(argv[1]);
        if (!job) throw std::runtime_error("cannot read the job file");
        const std::string out = argv[2];
        std::vector<std::unique_ptr<Variant>> vars;
        unsigned grid = 1;
        int compat = 0;
        uint64_t cell_cap = 1 << 16, stride_extra = 0, err_cap = 8;
        std::string line;
        while (std::getline(job, line)) {
            emu_unpoison(line.data(), line.size());
            std::istringstream ls(line);
            std::string key;
            if (!(ls >> key)) continue;
            emu_unpoison(key.data(), key.size());
            if (key == "grid") ls >> grid;
            else if (key == "compat") ls >> compat;
            else if (key == "cell_cap") ls >> cell_cap;
            else if (key == "stride_extra") ls >> stride_extra;
            else if (key == "err_cap") ls >> err_cap;
            else if (key == "variant") {
                auto V = std::make_unique<Variant>();
                std::string bnd, cfg, mu, md, is;
                ls >> bnd >> cfg >> V->n >> V->count >> V->first_id >> V->seed >> V->vid >> mu >> md >> is;
                if (!ls) throw std::runtime_error("bad variant line: " + line);
                for (std::string* t : {&bnd, &cfg, &mu, &md, &is}) emu_unpoison(t->data(), t->size());
                emu_unpoison(&V->n, sizeof(V->n));
                emu_unpoison(&V->count, sizeof(V->count));
                emu_unpoison(&V->first_id, sizeof(V->first_id));
                emu_unpoison(&V->seed, sizeof(V->seed));
                emu_unpoison(&V->vid, sizeof(V->vid));
                mbg::parse_model(slurp(bnd), slurp(cfg), V->pm);
                V->cm = mbg::compile_model(V->pm);
                const std::vector<float> fu = floats(mu), fd = floats(md), fi = floats(is);
                V->rt = mbg::make_rate_tables(V->cm, fu.empty() ? nullptr : fu.data(), fd.empty() ? nullptr : fd.data(),
                                              fi.empty() ? nullptr : fi.data());
                vars.push_back(std::move(V));
            } else {
                throw std::runtime_error("unknown job line: " + line);
            }
        }
        if (vars.empty()) throw std::runtime_error("no variant in the job");
#if !MBG_MULTI
        if (vars.size() != 1) throw std::runtime_error("a single-run shape takes one variant");
#endif
        // lanes' state in HBM: one column per thread of the grid (plus stride_extra spare columns)
        const uint64_t stride = (uint64_t)grid * MBG_BLOCK + stride_extra;
        std::vector<u64> lane_state;
#if MBG_STATE_HBM
        lane_state.assign((size_t)stride * (MBG_STATE_HBM == 2 ? 1 : MBG_STATE_HBM == 3 || MBG_WSUM_REG == 1 ? 2 : 3) * MBG_NW,
                          0x5A5A5A5A5A5A5A5Aull);
        emu_poison(lane_state.data(), lane_state.size() * 8);
#endif
        // error columns: the lanes' window lists (one column per grid thread) and their overflow count
        std::vector<u64> err_s, err_v, err_miss(1, 0);
#if MBG_ERRORS
        err_s.assign((size_t)err_cap * (grid * (uint64_t)MBG_BLOCK), 0x3C3C3C3C3C3C3C3Cull);
        err_v.assign((size_t)err_cap * (grid * (uint64_t)MBG_BLOCK), 0x3C3C3C3C3C3C3C3Cull);
        emu_poison(err_s.data(), err_s.size() * 8);
        emu_poison(err_v.data(), err_v.size() * 8);
#endif
        int wmax = 1;
        for (auto& Vp : vars) {
            Variant& V = *Vp;
            const mbg::CompiledModel& cm = V.cm;
            // the shape must be one the runtime could choose for this model
            if (cm.NW > MBG_NW) throw std::runtime_error("the model has more words than the shape");
            if ((MBG_RATE32 && !mbg::rates_fit_u32(V.rt)) || (MBG_WSMAD && !mbg::rates_fit_i31(V.rt)) ||
                (MBG_COMPACT && !mbg::compact_fits(cm)) || (cm.any_generic && !MBG_GENERIC) ||
                (cm.any_prog && !MBG_PROG) || cm.max_inputs > MBG_MAX_IN || (cm.K > 20) != (MBG_SPARSE != 0) ||
                cm.discrete != (MBG_DISCRETE != 0))
                throw std::runtime_error("the launch shape does not fit this model");
            V.blob = mbg::pack_blob(cm, V.rt, MBG_RATE32 != 0, MBG_COMPACT != 0);
            V.rs = mbg::run_scales(cm, V.n);
            wmax = std::max(wmax, V.rs.W);
            const uint64_t S = MBG_SPARSE ? 0 : V.rs.S;
            V.stage.assign((size_t)V.count * mbg::stage_row(MBG_NW), 0);
            emu_poison(V.stage.data(), V.stage.size() * 8);  // written by the initial-state pass
            V.prob.assign((size_t)V.rs.W * S, 0);
            V.th.assign((size_t)V.rs.W, 0);
            V.fin.assign((size_t)S, 0);
            V.fp.assign((size_t)V.count * MBG_NW, 0);
            emu_poison(V.fp.data(), V.fp.size() * 8);  // fixed-point records, written before read
            V.summ.assign(MBG_TRACE ? (size_t)V.count * 4 : 1, 0);
            V.counters.assign(3, 0);  // fixed points, transitions, work queue
            V.istate = V.rt.istate;
#if MBG_SPARSE
            V.cells.assign(cell_cap, ~0ull);
            V.cellv.assign(cell_cap, 0);
            V.cell_miss.assign(1, 0);
#endif
#if MBG_ERRORS
            V.prob2.assign((size_t)V.rs.W * S, 0);
            V.th2.assign((size_t)V.rs.W, 0);
            if (MBG_SPARSE) V.cellv2.assign(cell_cap, 0);
#endif
            mbg_desc& D = V.D;
            D.blob = reinterpret_cast<const char*>(V.blob.words.data());
            D.rtab = V.rt.rtab.empty() ? nullptr : reinterpret_cast<const u64*>(V.rt.rtab.data());
            D.istate = V.istate.data();
            D.stage = V.stage.data();
            D.prob = V.prob.data();
            D.th = V.th.data();
            D.fin = V.fin.data();
            D.fp_states = V.fp.data();
            D.fp_count = &V.counters[0];
            D.summ = V.summ.data();
            D.work_counter = &V.counters[2];
            D.flip_total = &V.counters[1];
            D.fp_cap = V.count;
            D.summ_cap = MBG_TRACE ? V.count : 0;
            D.nsteps = V.rs.nsteps;
            D.batch = V.count;
            D.batch_first = V.first_id;
            D.first_id = V.first_id;
            D.seed = V.seed;
            D.maxt = (double)cm.max_time;
            D.tick = (double)cm.time_tick;
            D.inv_tick = 1.0 / D.tick;
            D.rscale = std::ldexp(1.0, -V.rt.E);
            std::memcpy(&D.lay, &V.blob.lay, sizeof(D.lay));
            D.variant_id = V.vid;
            D.N = cm.N;
            D.K = cm.K;
            D.W = V.rs.W;
            D.FP = V.rs.FP;
            D.FT = V.rs.FT;
            D.sigma = V.rt.sigma;
            D.compat = compat;
#if MBG_SPARSE
            D.cells = V.cells.data();
            D.cell_vals = V.cellv.data();
            D.cell_miss = V.cell_miss.data();
            D.cell_mask = cell_cap - 1;
#endif
            D.lane_state = lane_state.empty() ? nullptr : lane_state.data();
            D.lane_stride = (u32)stride;
#if MBG_ERRORS
            D.prob2 = V.prob2.data();
            D.th2 = V.th2.data();
            D.cell_vals2 = V.cellv2.empty() ? nullptr : V.cellv2.data();
            D.err_s = err_s.data();
            D.err_v = err_v.data();
            D.err_miss = err_miss.data();
            D.sq_scale = V.rs.sq_scale;
            D.sq_scale_th = V.rs.sq_scale_th;
            D.FQ = V.rs.FQ;
            D.FQT = V.rs.FQT;
            D.err_cap = (u32)err_cap;
            D.err_stride = (u32)(grid * (uint64_t)MBG_BLOCK);
#endif
        }
        // everything the host prepared counts as written; what the kernels must write first
        // (staging rows, fixed-point records, lanes' state in HBM, window lists) does not
        for (auto& Vp : vars) {
            Variant& V = *Vp;
            auto clean = [](auto& vec) { emu_unpoison(vec.data(), vec.size() * sizeof(vec[0])); };
            clean(V.blob.words);
            clean(V.rt.rtab);
            clean(V.istate);
            clean(V.prob);
            clean(V.th);
            clean(V.fin);
            clean(V.summ);
            clean(V.counters);
            clean(V.cells);
            clean(V.cellv);
            clean(V.cell_miss);
            clean(V.prob2);
            clean(V.th2);
            clean(V.cellv2);
            emu_unpoison(&V.D, sizeof(V.D));
            emu_poison(V.stage.data(), V.stage.size() * 8);
            emu_poison(V.fp.data(), V.fp.size() * 8);
        }
        emu_unpoison(err_miss.data(), err_miss.size() * 8);
        emu_unpoison(&wmax, sizeof(wmax));
        emu_unpoison(&grid, sizeof(grid));
        const size_t th_cells = MBG_ERRORS ? 2 : 1;  // window entropies, then their squares
        const size_t lane_words =
            MBG_STATE_HBM == 1 ? 0 : (size_t)(MBG_STATE_HBM >= 2 ? 1 : MBG_WSUM_REG == 1 ? 2 : 3) * MBG_NW * MBG_BLOCK;
#if !MBG_MULTI
        Variant& V = *vars[0];
        mbg_desc& D = V.D;
        // initial states (static LDS: one block at a time)
        if (MBG_NW <= 128) {
            emu::launch({(unsigned)((V.count + 63) / 64), 1, 1}, {256, 1, 1}, 0, false, [&] { mbg_init_rates(D); });
        } else {
            const unsigned b = (unsigned)((V.count * MBG_NW + 255) / 256);
#if MBG_NW > 128
            emu::launch({b, 1, 1}, {256, 1, 1}, 0, false, [&] { mbg_init_state(D); });
#endif
            emu::launch({b, 1, 1}, {256, 1, 1}, 0, false, [&] { mbg_init_rates(D); });
            emu::launch({(unsigned)((V.count + 255) / 256), 1, 1}, {256, 1, 1}, 0, false, [&] { mbg_init_readouts(D); });
        }
        size_t smem_words = (MBG_TABLES_LDS ? V.blob.lay.bytes / 8 : 0) + lane_words;
        if (MBG_TH_LDS) smem_words += th_cells * (size_t)V.rs.W;
        if (MBG_CELLS_LDS) smem_words += ((size_t)V.rs.W + 1) * ((size_t)1 << V.cm.K);
        emu::launch({grid, 1, 1}, {MBG_BLOCK, 1, 1}, smem_words * 8, true, [&] { mbg_simulate(D); });
#else
        std::vector<mbg_desc> descs;
        uint64_t bx = 1;
        for (auto& Vp : vars) {
            descs.push_back(Vp->D);
            bx = std::max<uint64_t>(bx, (Vp->count + 63) / 64);
        }
        const mbg_desc* dp = descs.data();
        emu_unpoison(descs.data(), descs.size() * sizeof(mbg_desc));
        emu_unpoison(&dp, sizeof(dp));
        emu::launch({(unsigned)bx, (unsigned)descs.size(), 1}, {256, 1, 1}, 0, false, [&] { mbg_init_rates_multi(dp); });
        // tables in LDS: the class's largest blob, then a 16-byte flag, before the lanes' state
        int tab_words = 0;
        if (MBG_TABLES_LDS)
            for (auto& Vp : vars) tab_words = std::max(tab_words, (int)(Vp->blob.lay.bytes / 8));
        const size_t smem_words = (MBG_TABLES_LDS ? (size_t)tab_words + 2 : 0) + lane_words +
                                  (MBG_TH_LDS ? (size_t)(MBG_BLOCK / 64) * th_cells * (size_t)wmax : 0);
        const int nv = (int)descs.size();
        // blocks start at variants spread over the class (block b at variant b mod nv)
        std::vector<int> starts(grid);
        for (unsigned b = 0; b < grid; b++) starts[b] = (int)(b % (unsigned)nv);
        const int* sp = MBG_TABLES_LDS ? starts.data() : nullptr;
        emu_unpoison(&nv, sizeof(nv));
        emu_unpoison(&tab_words, sizeof(tab_words));
        emu_unpoison(&sp, sizeof(sp));
        emu_unpoison(starts.data(), starts.size() * sizeof(int));
        emu::launch({grid, 1, 1}, {MBG_BLOCK, 1, 1}, smem_words * 8, true,
                    [&] { mbg_simulate_multi(dp, nv, wmax, nullptr, tab_words, sp); });
#endif
        for (size_t v = 0; v < vars.size(); v++) {
            Variant& X = *vars[v];
            const std::string p = out + "/v" + std::to_string(v) + "_";
            const uint64_t nfp = X.counters[0];
            if (nfp > X.count) throw std::runtime_error("fixed-point count beyond the trajectories");
            const uint64_t info[11] = {X.counters[1], nfp, (uint64_t)X.rs.W, (uint64_t)X.cm.K, (uint64_t)X.rs.FP,
                                       (uint64_t)X.rs.FT, (uint64_t)MBG_NW, X.cell_miss.empty() ? 0 : X.cell_miss[0],
                                       err_miss[0], (uint64_t)X.rs.FQ, (uint64_t)X.rs.FQT};
            dump(p + "info.u64", info, sizeof(info));
            dump(p + "prob.u64", X.prob.data(), X.prob.size() * 8);
            dump(p + "th.u64", X.th.data(), X.th.size() * 8);
            dump(p + "fin.u64", X.fin.data(), X.fin.size() * 8);
            dump(p + "fp.u64", X.fp.data(), (size_t)nfp * MBG_NW * 8);
            dump(p + "summ.u64", X.summ.data(), MBG_TRACE ? X.summ.size() * 8 : 0);
            dump(p + "cells.u64", X.cells.data(), X.cells.size() * 8);
            dump(p + "cellv.u64", X.cellv.data(), X.cellv.size() * 8);
            dump(p + "prob2.u64", X.prob2.data(), X.prob2.size() * 8);
            dump(p + "th2.u64", X.th2.data(), X.th2.size() * 8);
            dump(p + "cellv2.u64", X.cellv2.data(), X.cellv2.size() * 8);
        }
    } catch (const std::exception& e) {
        std::cerr << "emu_run: " << e.what() << "\n";
        return 1;
    }
    return 0;
}
