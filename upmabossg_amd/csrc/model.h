// Compiled model: the data the trajectory kernel reads, built from a ParsedModel.
//
// The reference turns each node into generated CUDA (src/generator.cpp:84-122):
//   fast path  rate = (x_i == logic_i(x)) ? 0 : (x_i ? D_i : U_i)   (generator.cpp:93-107)
//   generic    rate = x_i ? eval(rate_down) : eval(rate_up)          (generator.cpp:108-119)
// Here each node becomes a truth table over its key = x_i | (input bits << 1):
//   fast path  one logic bit per key, plus the FP32 constants U_i, D_i
//   generic    one FP32 rate per key
// so that every lane of a wave evaluates any node with the same instructions.  A fast node
// that reads more than kMaxInputsFast nodes (the ring models of data/generate-synth.py read
// 2 sqrt(N) + 2) gets a postfix logic program instead, run by the same loop on every lane.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "front.h"

namespace mbg {

// Layout shared with struct mbg_node in device/mbg_kernel.hip (16 bytes).
struct NodeRec {
    uint32_t in_off;
    uint32_t dep_off;
    uint32_t tab_off;
    uint16_t dep_cnt;
    uint8_t in_cnt;
    uint8_t flags;  // bit 7: generic; bit 6: listed among its own dependents; bits 0-5: readout index + 1
};
static_assert(sizeof(NodeRec) == 16, "NodeRec must match the device layout");

constexpr int kMaxInputsFast = 20;     // LUT of 2^21 bits; wider fast nodes run a logic program
constexpr uint8_t kProgInputs = 255;   // NodeRec::in_cnt of a node whose logic is a program
constexpr int kMaxProgDepth = 64;      // bit-stack depth of a logic program
constexpr int kMaxInputsGeneric = 16;  // table of 2^17 rates
constexpr int kMaxReadouts = 20;       // dense window tables (the reference's limit, src/main.cpp:142-146)
constexpr int kMaxSparseReadouts = 63; // sparse window tables (hash of window and projected state)
constexpr const char* kReadoutLimitMessage = "This executable supports a maximum of 63 non-internal nodes.";

struct CompiledModel {
    std::vector<std::string> names;
    int N = 0, NW = 0, K = 0;
    std::vector<int> readouts;        // non-internal nodes in node order
    std::vector<uint8_t> internal;
    bool discrete = false;
    float max_time = 10.f, time_tick = 0.1f;
    double sample_count = 0, seed = 0;

    std::vector<NodeRec> nodes;
    std::vector<uint32_t> inputs, deps, lut;
    // logic programs of fast nodes with more than kMaxInputsFast inputs (in_cnt = kProgInputs,
    // tab_off = offset of [length, ops...]); op format in model.cpp
    std::vector<uint32_t> prog;
    bool any_prog = false;
    std::vector<uint8_t> generic;
    std::vector<uint64_t> gen_mask;   // per 64-node word
    std::vector<float> up, down;      // fast nodes, FP32
    std::vector<float> gtab;          // generic nodes: FP32 rate per key, at NodeRec::tab_off
    std::vector<float> istate;
    bool any_generic = false;
    int max_inputs = 0;               // largest in_cnt of a truth-table node
};

// Entropy term of a fixed-point rate: rint(r * log2(r) * 2^sigma), 0 for r = 0.
uint64_t entropy_term(int sigma, uint64_t r);

// Per-run (per-variant) numeric tables in the kernel's fixed-point units.
struct RateTables {
    int E = 0;                        // rate_fixed = rate * 2^E
    std::vector<uint64_t> rates;      // [U, D] per node
    std::vector<uint64_t> rtab;       // generic tables
    int sigma = 62;                   // entropy term = rint(r * log2(r) * 2^sigma)
    std::vector<uint64_t> terms;      // per readout: entropy term of its U and D
    std::vector<float> istate;
};

CompiledModel compile_model(const ParsedModel& pm);

// Applies optional per-node multipliers on rate_up / rate_down and an optional istate
// override (all may be null) and converts to fixed point.
RateTables make_rate_tables(const CompiledModel& cm, const float* mult_up, const float* mult_down,
                            const float* istate);

// FP32 rate of node i at a state (NW words), from the compiled tables: the truth table or
// logic program of a fast node with U_i / D_i, or the generic rate table (no multipliers).
float node_rate(const CompiledModel& cm, int i, const uint64_t* state);

int ceil_log2_u64(uint64_t v);
int ceil_log2_real(double q);

}  // namespace mbg
