// Kernel-facing tables of a model, built on the host without any HIP call: the packed table
// blob the trajectory kernel reads (struct mbg_layout / mbg_node in device/mbg_kernel.hip),
// the rate-width tests that choose its layout, and the per-run window count and fixed-point
// exponents.  Shared by the runtime (runtime.cpp) and by anything else that lays out the
// kernel's inputs on the host.
#pragma once

#include <cstdint>
#include <vector>

#include "model.h"

namespace mbg {

// Byte offsets of the model tables in one blob; identical to struct mbg_layout in
// device/mbg_kernel.hip.
struct Layout {
    uint32_t nodes, rates, deps, inputs, lut, readouts, terms, gen_mask, bytes, prog;
    uint32_t onehot;  // u32[16][16] identity: row g holds the factors [i == g] of a rate-sum update
};

// Model tables packed into one blob (16-byte aligned segments).
struct Blob {
    Layout lay{};
    std::vector<uint64_t> words;
};

// Every fixed-point rate fits 32 bits (the rate table is u32, MBG_RATE32).
bool rates_fit_u32(const RateTables& rt);
// Every rate (fast and generic tables) below 2^31: any change of one rate fits an i32, so the
// register rate sums can be updated with one 32 x 32 + 64 multiply-add each (MBG_WSMAD).
bool rates_fit_i31(const RateTables& rt);
// The compact table layout (8-byte node records, 16-bit dependents, MBG_COMPACT) fits when the
// model has no generic rate tables and its tables stay within the packed field widths.
bool compact_fits(const CompiledModel& cm);

Blob pack_blob(const CompiledModel& cm, const RateTables& rt, bool rate32, bool compact);

// 64-bit words of a staged initial state (x, act and word sums, then the readout part:
// projected state, readout rate sum, entropy-term sum, active-readout count), as STAGE_ROW in
// the kernel.
inline uint64_t stage_row(int nw) { return 3ull * (uint64_t)nw + 4ull; }

// Window count, fixed-point exponents and discrete step count of a run of n trajectories, and
// for the error columns the exponents FQ, FQT of the sums of squared per-trajectory window
// totals and the scales turning a total into the trajectory's fraction of the window
// (sq_scale) or its window-average TH (sq_scale_th).
struct RunScales {
    int W = 0, FP = 0, FT = 0;
    uint64_t S = 0, nsteps = 0;
    int FQ = 0, FQT = 0;
    double sq_scale = 1.0, sq_scale_th = 1.0;
};
RunScales run_scales(const CompiledModel& cm, uint64_t n);

}  // namespace mbg
