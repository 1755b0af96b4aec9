/*
 * CPU oracle for the MaBoSS-semantics trajectory kernel and statistics stage.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or the timed CPU baseline).  The product (libmbg.so, the MaBoSSG CLI) never
 * links or calls it.
 *
 * The oracle restates the reference algorithm of MaBoSSG with its own parser
 * and expression interpreter (independent of the product's front end):
 *   - language:  src/parser/scanner.ll:16-93, src/parser/parser.yy:74-158,
 *                src/parser/driver.cpp:7-99, src/parser/parse_types.cpp:3-22
 *   - node rates: src/generator.cpp:57-122, src/parser/expressions.cpp:10-231
 *   - step loop: src/jit_kernels/simulation.cu:9-164 (Gillespie, flip scan)
 *   - stats:     src/jit_kernels/window_average_small.cu:50-136,
 *                src/jit_kernels/final_states.cu:9-21,
 *                src/statistics/fixed_states.cu:84-187
 * with the documented deviations of this project (FP64 clock, Philox4x32-10
 * keyed by trajectory id, exact fixed-point rate sums, fixed-point window
 * accumulators) so that the GPU kernel can be replayed bit-exactly.
 *
 * Two extensions beyond the reference's limits follow the same semantics: any
 * number of inputs per fast node (expressions are interpreted, not tabulated),
 * and more than 20 readouts, where the window cells and final states are kept
 * in per-thread maps instead of dense tables (pinned against the dense path by
 * marginalising a 24-readout run onto 20 readouts, tests/test_sparse_oracle.py).
 */
#pragma once
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_model orc_model;

/* Parse a .bnd/.cfg pair given as text.  Returns NULL on error (message in err). */
orc_model* orc_load_text(const char* bnd, const char* cfg, char* err, size_t errlen);
void orc_free(orc_model* m);

int orc_num_nodes(const orc_model* m);
const char* orc_node_name(const orc_model* m, int i);
float orc_node_istate(const orc_model* m, int i);
int orc_node_is_internal(const orc_model* m, int i);
int orc_node_is_fast(const orc_model* m, int i);
/* Run constants (discrete_time, sample_count, max_time, time_tick, seed_pseudorandom,
 * or any constant of the cfg); returns nan when absent. */
double orc_constant(const orc_model* m, const char* name);
/* Value of a $variable; nan when absent. */
double orc_variable(const orc_model* m, const char* name);

/* Per-variant parameter overrides (multipliers on up/down rates, istate). Any pointer may be NULL. */
int orc_set_variant(orc_model* m, const float* mult_up, const float* mult_down, const float* istate);

/* Fixed-point rate scale exponent E: rate_fixed = rate * 2^E. */
int orc_rate_exponent(const orc_model* m);
/* FP32 rate of node i in state x (bit i of word i/64); the value the reference's generated
 * <node>_rate function returns (src/generator.cpp:84-122). */
float orc_node_rate_f32(const orc_model* m, int i, const uint64_t* x);
/* Fixed-point rate of node i in state x. */
uint64_t orc_node_rate_fixed(const orc_model* m, int i, const uint64_t* x);

typedef struct orc_params {
    uint64_t n_trajectories;   /* normalisation count n (sample_count) */
    uint64_t first_id;         /* simulate ids [first_id, first_id + count) */
    uint64_t count;
    uint64_t seed;
    uint32_t variant_id;       /* Philox counter word 3 */
    int threads;               /* worker threads (>= 1) */
    int record_fixed;          /* collect fixed-point states */
    int record_summaries;      /* fill per-trajectory summaries */
    int compat_batch_resume;   /* 1: no flip after every 99th step, as the reference's 100-record
                                  launches do (src/jit_kernels/simulation.cu:104-147) */
} orc_params;

typedef struct orc_summary {
    uint64_t flips;            /* transitions taken */
    uint64_t flip_hash;        /* FNV-1a over flipped node indices */
    double final_time;
    uint32_t status;           /* 1 = FINISHED, 2 = FIXED_POINT */
    uint32_t pad;
} orc_summary;

typedef struct orc_result orc_result;

orc_result* orc_run(orc_model* m, const orc_params* p, char* err, size_t errlen);
void orc_result_free(orc_result* r);

int orc_windows(const orc_result* r);
int orc_readouts(const orc_result* r);
int orc_prob_exponent(const orc_result* r);  /* F_P: cell = sum(rint(slice * 2^F_P)) */
int orc_th_exponent(const orc_result* r);    /* F_TH */
/* W*S cells (window-major), W TH cells, S final-state counts. */
const uint64_t* orc_prob_table(const orc_result* r);
const uint64_t* orc_th_table(const orc_result* r);
const uint64_t* orc_final_table(const orc_result* r);
/* The non-zero cells, window-major with states ascending (every run; the only form of the
 * window table and final states above 20 readouts, where prob and final tables are empty). */
uint64_t orc_cell_count(const orc_result* r);
const uint32_t* orc_cell_windows(const orc_result* r);
const uint64_t* orc_cell_states(const orc_result* r);
const uint64_t* orc_cell_values(const orc_result* r);
/* Error columns: sums over trajectories of squared per-trajectory window totals, W*S cells
 * (window-major) and W TH cells, in units of 2^-FQ (resp. 2^-FQT) of the squared fraction of
 * the window (resp. squared window-average TH); orc_cell_squares follows orc_cell_values. */
const uint64_t* orc_prob_sq_table(const orc_result* r);
const uint64_t* orc_th_sq_table(const orc_result* r);
int orc_sq_exponent(const orc_result* r);
int orc_th_sq_exponent(const orc_result* r);
const uint64_t* orc_cell_squares(const orc_result* r);
uint64_t orc_final_cell_count(const orc_result* r);
const uint64_t* orc_final_cell_states(const orc_result* r);
const uint64_t* orc_final_cell_counts(const orc_result* r);
uint64_t orc_total_flips(const orc_result* r);
double orc_elapsed_seconds(const orc_result* r);
/* Fixed points: sorted ascending by the reference order (most significant word first). */
uint64_t orc_fixed_count(const orc_result* r);                    /* distinct states */
int orc_state_words(const orc_result* r);                         /* 64-bit words per state */
const uint64_t* orc_fixed_states(const orc_result* r);            /* count * words */
const uint64_t* orc_fixed_counts(const orc_result* r);
const orc_summary* orc_summaries(const orc_result* r);            /* count entries or NULL */

/* Building blocks, exported for known-answer tests. */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
double orc_ln(double x);
double orc_log2(double x);
double orc_neglog_u32(uint32_t x);

#ifdef __cplusplus
}
#endif
