// Kernel-facing tables of a model (pack.h): blob packing and run scales, host only.
#include "pack.h"

#include <cmath>
#include <cstring>

namespace mbg {

bool rates_fit_u32(const RateTables& rt) {
    for (uint64_t r : rt.rates)
        if (r >> 32) return false;
    return true;
}

bool rates_fit_i31(const RateTables& rt) {
    for (uint64_t r : rt.rates)
        if (r >> 31) return false;
    for (uint64_t r : rt.rtab)
        if (r >> 31) return false;
    return true;
}

bool compact_fits(const CompiledModel& cm) {
    if (cm.any_generic || cm.any_prog || cm.N > 4096 || cm.K > 31 || cm.max_inputs > 15) return false;
    if (cm.inputs.size() + 32 >= (1u << 14) || cm.deps.size() >= (1u << 14)) return false;
    for (const NodeRec& r : cm.nodes)
        if (r.dep_cnt > 255 || (uint64_t)r.tab_off + (2ull << r.in_cnt) > (1ull << 16)) return false;
    return true;
}

Blob pack_blob(const CompiledModel& cm, const RateTables& rt, bool rate32, bool compact) {
    Blob b;
    std::vector<uint8_t> bytes;
    auto seg = [&](const void* p, size_t n) -> uint32_t {
        size_t off = (bytes.size() + 15) & ~(size_t)15;
        bytes.resize(off + n);
        if (n) std::memcpy(bytes.data() + off, p, n);
        return (uint32_t)off;
    };
    std::vector<uint16_t> in16(cm.inputs.begin(), cm.inputs.end());
    in16.resize(in16.size() + 32, 0);  // the kernel's unrolled gathers may read past the last node
    std::vector<uint32_t> ro(cm.readouts.begin(), cm.readouts.end());
    if (compact) {
        // bits 0-13 in_off, 14-27 dep_off, 28-43 tab_off, 44-51 dep_cnt, 52-55 in_cnt,
        // 56 listed among its own dependents, 57-61 readout index + 1
        std::vector<uint64_t> nodes(cm.nodes.size());
        for (size_t i = 0; i < nodes.size(); i++) {
            const NodeRec& r = cm.nodes[i];
            nodes[i] = (uint64_t)r.in_off | (uint64_t)r.dep_off << 14 | (uint64_t)r.tab_off << 28 |
                       (uint64_t)r.dep_cnt << 44 | (uint64_t)r.in_cnt << 52 | (uint64_t)((r.flags >> 6) & 1u) << 56 |
                       (uint64_t)(r.flags & 0x3fu) << 57;
        }
        b.lay.nodes = seg(nodes.data(), nodes.size() * 8);
    } else {
        b.lay.nodes = seg(cm.nodes.data(), cm.nodes.size() * sizeof(NodeRec));
    }
    if (rate32) {
        std::vector<uint32_t> r32(rt.rates.begin(), rt.rates.end());
        b.lay.rates = seg(r32.data(), r32.size() * 4);
    } else {
        b.lay.rates = seg(rt.rates.data(), rt.rates.size() * 8);
    }
    if (compact) {
        // node id (12 bits) | key bit position (4 bits)
        std::vector<uint16_t> d16(cm.deps.size());
        for (size_t i = 0; i < d16.size(); i++)
            d16[i] = (uint16_t)((cm.deps[i] & 0xfffu) | ((cm.deps[i] >> 24) & 0xfu) << 12);
        b.lay.deps = seg(d16.data(), d16.size() * 2);
    } else {
        b.lay.deps = seg(cm.deps.data(), cm.deps.size() * 4);
    }
    b.lay.inputs = seg(in16.data(), in16.size() * 2);
    b.lay.lut = seg(cm.lut.data(), cm.lut.size() * 4);
    b.lay.readouts = seg(ro.data(), ro.size() * 4);
    b.lay.terms = seg(rt.terms.data(), rt.terms.size() * 8);
    b.lay.gen_mask = seg(cm.gen_mask.data(), cm.gen_mask.size() * 8);
    b.lay.prog = seg(cm.prog.data(), cm.prog.size() * 4);
    // factors [i == g] of the register rate-sum updates, read as four 16-byte rows
    uint32_t onehot[16 * 16] = {};
    for (int g = 0; g < 16; g++) onehot[g * 16 + g] = 1u;
    b.lay.onehot = seg(onehot, sizeof(onehot));
    bytes.resize((bytes.size() + 15) & ~(size_t)15);
    b.lay.bytes = (uint32_t)bytes.size();
    b.words.resize(bytes.size() / 8);
    std::memcpy(b.words.data(), bytes.data(), bytes.size());
    return b;
}


RunScales run_scales(const CompiledModel& cm, uint64_t n) {
    RunScales r;
    r.W = (int)std::ceil(cm.max_time / cm.time_tick);  // window_average_small.cpp:26
    r.S = 1ull << cm.K;
    if (cm.discrete) {
        r.FP = 0;
        r.FT = 59 - ceil_log2_u64(n);
    } else {
        const int eq = ceil_log2_real((double)n * (double)cm.time_tick);
        r.FP = 60 - eq;
        r.FT = 57 - eq;
    }
    const float ratio = cm.max_time / cm.time_tick;
    r.nsteps = ratio >= 2.f ? (uint64_t)std::floor(ratio) : 1;
    // a window fraction is at most 1 and a window-average TH below 64 (at most 6 bits of
    // entropy over 63 readouts): n squares stay below 2^62
    r.FQ = 62 - ceil_log2_u64(n);
    r.FQT = 56 - ceil_log2_u64(n);
    const double tick = (double)cm.time_tick;
    r.sq_scale = cm.discrete ? 1.0 : std::ldexp(1.0, -r.FP) / tick;
    r.sq_scale_th = cm.discrete ? std::ldexp(1.0, -r.FT) : std::ldexp(1.0, -r.FT) / tick;
    return r;
}

}  // namespace mbg
