// Output formats of the reference: <prefix>_probtraj.csv, <prefix>_fp.csv and the stdout dump.
#pragma once

#include <cstdint>
#include <ostream>
#include <string>
#include <vector>

#include "../../include/mbg.h"
#include "model.h"

namespace mbg {

// Host allocator for page-locked memory (fast device-to-host copies of large tables).
template <class T>
struct PinnedAllocator {
    using value_type = T;
    PinnedAllocator() = default;
    template <class U>
    PinnedAllocator(const PinnedAllocator<U>&) {}
    T* allocate(size_t n);
    void deallocate(T* p, size_t) noexcept;
    bool operator==(const PinnedAllocator&) const { return true; }
    bool operator!=(const PinnedAllocator&) const { return false; }
};
using PinnedU64 = std::vector<uint64_t, PinnedAllocator<uint64_t>>;

struct Results {
    const mbg_run_info* info;
    const PinnedU64* prob;                      // [W][2^K] fixed point (counts when discrete)
    const std::vector<uint64_t>* th;            // [W]
    const std::vector<uint64_t>* fin;           // [2^K]
    const std::vector<uint64_t>* fixed_states;  // [F][state_words], reference order
    const std::vector<uint64_t>* fixed_counts;  // [F]
    // More than 20 readouts (prob and fin empty): the non-zero window cells, window-major with
    // states ascending, and the final states, ascending.
    const std::vector<uint32_t>* cell_w = nullptr;
    const std::vector<uint64_t>* cell_s = nullptr;
    const std::vector<uint64_t>* cell_v = nullptr;
    const std::vector<uint64_t>* fin_s = nullptr;
    const std::vector<uint64_t>* fin_v = nullptr;
    // Error columns (null when the run did not accumulate them: the columns print 0): sums of
    // squared per-trajectory window totals, beside prob (or cell_v) and th.
    const std::vector<uint64_t>* prob2 = nullptr;
    const std::vector<uint64_t>* cell_v2 = nullptr;
    const std::vector<uint64_t>* th2 = nullptr;
};

// Normalised probability of a raw cell value.
double cell_prob(const mbg_run_info& I, uint64_t v);

// Normalised probability of cell (window-major) and window transition entropy
// (src/statistics/window_average_small.cpp:126-140, :222-223).
double window_prob(const Results& r, size_t cell);
double window_th(const Results& r, size_t w);

// Standard error of the mean over the n trajectories of a per-trajectory window quantity with
// mean `mean` and sum of squares raw_sq (units 2^-exponent): sqrt((E[x^2] - mean^2) / (n - 1)),
// 0 when n < 2 or the difference rounds below 0.  The ErrorProba and ErrorTH columns.
double standard_error(const mbg_run_info& I, double mean, uint64_t raw_sq, int exponent);
double window_prob_error(const Results& r, size_t cell);
double window_th_error(const Results& r, size_t w);

// " -- "-joined names of the ON nodes, or "<nil>" (src/state.cpp:29-48).
std::string state_string(const CompiledModel& cm, const uint64_t* words);
std::string projected_string(const CompiledModel& cm, uint64_t s);

bool write_probtraj_csv(const Results& r, const CompiledModel& cm, const std::string& path);
bool write_fp_csv(const Results& r, const CompiledModel& cm, const std::string& path);
void write_text(const Results& r, const CompiledModel& cm, std::ostream& os);

}  // namespace mbg
