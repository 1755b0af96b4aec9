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
    float time, th, entropy;
    std::vector<std::pair<uint64_t, float>> probs;  // non-zero states ascending
};

WindowRow window_row(const Results& r, size_t w) {
    WindowRow row;
    row.time = (float)w * r.info->time_tick;
    row.th = (float)window_th(r, w);
    row.entropy = 0.f;
    auto add = [&](uint64_t s, uint64_t raw) {
        const float p = (float)cell_prob(*r.info, raw);
        if (p == 0.f) return;
        row.entropy += -std::log2(p) * p;
        row.probs.emplace_back(s, p);
    };
    if (r.cell_w) {
        const auto lo = std::lower_bound(r.cell_w->begin(), r.cell_w->end(), (uint32_t)w);
        const auto hi = std::upper_bound(lo, r.cell_w->end(), (uint32_t)w);
        for (size_t i = (size_t)(lo - r.cell_w->begin()); i < (size_t)(hi - r.cell_w->begin()); i++)
            add((*r.cell_s)[i], (*r.cell_v)[i]);
    } else {
        const uint64_t S = 1ull << r.info->readouts;
        for (uint64_t s = 0; s < S; s++) add(s, (*r.prob)[w * S + s]);
    }
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
    for (const WindowRow& row : rows) {
        ofs << row.time << "\t" << row.th << "\t" << 0.f << "\t" << row.entropy << "\t" << 0.f;
        for (const auto& sp : row.probs) ofs << "\t" << projected_string(cm, sp.first) << "\t" << sp.second << "\t" << 0.f;
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
        for (const auto& sp : row.probs) os << sp.second << " " << projected_string(cm, sp.first) << std::endl;
    }
}

}  // namespace mbg
