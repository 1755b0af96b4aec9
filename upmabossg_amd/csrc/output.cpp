// Writers for the reference's output formats.  Values are printed as FP32 with the default
// ostream formatting (6 significant digits), tab separated, one std::endl per row, as
// src/statistics/window_average_small.cpp:142-250, src/statistics/fixed_states.cu:141-187 and
// src/statistics/final_states.cpp:53-66 print them.
#include "output.h"

#include <algorithm>
#include <cmath>
#include <fstream>

namespace mbg {

double cell_prob(const mbg_run_info& I, uint64_t raw) {
    const double v = (double)raw;
    if (I.discrete) return v / (double)I.n_trajectories;
    return std::ldexp(v, -I.prob_exponent) / ((double)I.n_trajectories * (double)I.time_tick);
}

double window_prob(const Results& r, size_t cell) { return cell_prob(*r.info, (*r.prob)[cell]); }

double window_th(const Results& r, size_t w) {
    const mbg_run_info& I = *r.info;
    double v = std::ldexp((double)(*r.th)[w], -I.th_exponent) / (double)I.n_trajectories;
    if (!I.discrete) v /= (double)I.time_tick;
    return v;
}

double standard_error(const mbg_run_info& I, double mean, uint64_t raw_sq, int exponent) {
    const double n = (double)I.n_trajectories;
    if (n < 2.0) return 0.0;
    const double var = std::ldexp((double)raw_sq, -exponent) / n - mean * mean;
    return var > 0.0 ? std::sqrt(var / (n - 1.0)) : 0.0;
}

double window_prob_error(const Results& r, size_t cell) {
    if (!r.prob2) return 0.0;
    return standard_error(*r.info, window_prob(r, cell), (*r.prob2)[cell], r.info->sq_exponent);
}

double window_th_error(const Results& r, size_t w) {
    if (!r.th2) return 0.0;
    return standard_error(*r.info, window_th(r, w), (*r.th2)[w], r.info->th_sq_exponent);
}

std::string state_string(const CompiledModel& cm, const uint64_t* words) {
    std::string s;
    for (int i = 0; i < cm.N; i++)
        if ((words[i >> 6] >> (i & 63)) & 1ull) {
            if (!s.empty()) s += " -- ";
            s += cm.names[i];
        }
    return s.empty() ? "<nil>" : s;
}

std::string projected_string(const CompiledModel& cm, uint64_t st) {
    std::string s;
    for (int j = 0; j < cm.K; j++)
        if ((st >> j) & 1ull) {
            if (!s.empty()) s += " -- ";
            s += cm.names[cm.readouts[j]];
        }
    return s.empty() ? "<nil>" : s;
}

namespace {

struct WindowRow {
    float time, th, th_err, entropy, hd0;
    struct Cell {
        uint64_t state;
        float p, err;
    };
    std::vector<Cell> probs;  // non-zero states ascending
};

// One window of the probtraj output.  HD=0 is the probability that the state is at Hamming
// distance 0 from the reference state; the language has no reference state, so every state is
// at distance 0 and the column is the window's total probability.
WindowRow window_row(const Results& r, size_t w) {
    const mbg_run_info& I = *r.info;
    WindowRow row;
    row.time = (float)w * I.time_tick;
    const double th = window_th(r, w);
    row.th = (float)th;
    row.th_err = r.th2 ? (float)standard_error(I, th, (*r.th2)[w], I.th_sq_exponent) : 0.f;
    row.entropy = 0.f;
    double total = 0.0;
    auto add = [&](uint64_t s, uint64_t raw, const uint64_t* raw_sq) {
        const double pd = cell_prob(I, raw);
        const float p = (float)pd;
        if (p == 0.f) return;
        row.entropy += -std::log2(p) * p;
        total += pd;
        row.probs.push_back({s, p, raw_sq ? (float)standard_error(I, pd, *raw_sq, I.sq_exponent) : 0.f});
    };
    if (r.cell_w) {
        const auto lo = std::lower_bound(r.cell_w->begin(), r.cell_w->end(), (uint32_t)w);
        const auto hi = std::upper_bound(lo, r.cell_w->end(), (uint32_t)w);
        for (size_t i = (size_t)(lo - r.cell_w->begin()); i < (size_t)(hi - r.cell_w->begin()); i++)
            add((*r.cell_s)[i], (*r.cell_v)[i], r.cell_v2 ? &(*r.cell_v2)[i] : nullptr);
    } else {
        const uint64_t S = 1ull << I.readouts;
        for (uint64_t s = 0; s < S; s++) add(s, (*r.prob)[w * S + s], r.prob2 ? &(*r.prob2)[w * S + s] : nullptr);
    }
    row.hd0 = r.th2 ? (float)total : 0.f;
    return row;
}

}  // namespace

bool write_probtraj_csv(const Results& r, const CompiledModel& cm, const std::string& path) {
    std::ofstream ofs(path);
    if (!ofs) return false;
    const size_t W = (size_t)r.info->windows;
    std::vector<WindowRow> rows;
    size_t max_states = 0;
    for (size_t w = 0; w < W; w++) {
        rows.push_back(window_row(r, w));
        max_states = std::max(max_states, rows.back().probs.size());
    }
    ofs << "Time\tTH\tErrorTH\tH\tHD=0";
    for (size_t i = 0; i < max_states; i++) ofs << "\tState\tProba\tErrorProba";
    ofs << std::endl;
    // ErrorTH, HD=0 and ErrorProba are 0 when the run did not accumulate the error columns, as
    // the reference always writes them (window_average_small.cpp:235,245)
    for (const WindowRow& row : rows) {
        ofs << row.time << "\t" << row.th << "\t" << row.th_err << "\t" << row.entropy << "\t" << row.hd0;
        for (const auto& c : row.probs) ofs << "\t" << projected_string(cm, c.state) << "\t" << c.p << "\t" << c.err;
        ofs << std::endl;
    }
    return (bool)ofs;
}

bool write_fp_csv(const Results& r, const CompiledModel& cm, const std::string& path) {
    std::ofstream ofs(path);
    if (!ofs) return false;
    const size_t F = r.fixed_counts->size(), NW = (size_t)cm.NW;
    ofs << "Fixed Points (" << F << ")" << std::endl;
    ofs << "FP\tProba\tState";
    for (const auto& n : cm.names) ofs << "\t" << n;
    ofs << std::endl;
    for (size_t i = 0; i < F; i++) {
        const uint64_t* st = r.fixed_states->data() + i * NW;
        ofs << "#" << i << "\t" << (float)(*r.fixed_counts)[i] / (float)r.info->n_trajectories << "\t"
            << state_string(cm, st);
        for (int k = 0; k < cm.N; k++) ofs << "\t" << (int)((st[k >> 6] >> (k & 63)) & 1ull);
        ofs << std::endl;
    }
    return (bool)ofs;
}

void write_text(const Results& r, const CompiledModel& cm, std::ostream& os) {
    const float n = (float)r.info->n_trajectories;
    os << "final points:" << std::endl;
    if (r.fin_s) {
        for (size_t i = 0; i < r.fin_s->size(); i++)
            os << (float)(*r.fin_v)[i] / n << " " << projected_string(cm, (*r.fin_s)[i]) << std::endl;
    } else {
        for (size_t s = 0; s < r.fin->size(); s++)
            if ((*r.fin)[s]) os << (float)(*r.fin)[s] / n << " " << projected_string(cm, s) << std::endl;
    }
    os << "fixed points:" << std::endl;
    for (size_t i = 0; i < r.fixed_counts->size(); i++)
        os << (float)(*r.fixed_counts)[i] / n << " " << state_string(cm, r.fixed_states->data() + i * cm.NW) << std::endl;
    const float tick = r.info->time_tick;
    for (size_t w = 0; w < (size_t)r.info->windows; w++) {
        const WindowRow row = window_row(r, w);
        os << "window (" << (float)w * tick << ", " << (float)(w + 1) * tick << "]" << std::endl;
        os << "entropy: " << row.entropy << std::endl;
        os << "transition entropy: " << row.th << std::endl;
        for (const auto& c : row.probs) os << c.p << " " << projected_string(cm, c.state) << std::endl;
    }
}

}  // namespace mbg
