/*
 * mbg.h - C ABI of the MI355X (gfx950) MaBoSS-semantics trajectory library, libmbg.so.
 *
 * Plain C types only (no C++ or torch types cross it).  Every function returns an
 * mbg_status; errors never exit the process (the reference calls exit() from
 * CU_CHECK/CUDA_CHECK, src/utils.cpp:3-29).  A context is not thread-safe; use one
 * per host thread (as the reference, which runs on one host thread).  Device memory
 * is owned by the library; host output buffers are owned by the caller.
 *
 * Each entry point names the reference interface it replaces.
 */
#ifndef MBG_H
#define MBG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int mbg_status;
#define MBG_OK 0
#define MBG_ERR_PARSE 1      /* .bnd/.cfg syntax or semantic error */
#define MBG_ERR_LIMIT 2      /* model exceeds a limit (e.g. > 20 non-internal nodes) */
#define MBG_ERR_ARG 3        /* invalid argument or call order */
#define MBG_ERR_HIP 4        /* HIP runtime error (no device, out of memory, launch failure) */
#define MBG_ERR_COMPILE 5    /* hipRTC compilation failed */
#define MBG_ERR_NOT_BUILT 6  /* no result available yet */

typedef struct mbg_model mbg_model;
typedef struct mbg_ctx mbg_ctx;

/* ---------------------------------------------------------------- model (host only) */

/* Parses a .bnd text then a .cfg text.  Replaces driver::parse (src/parser/driver.cpp:16-31)
 * and the Flex/Bison front end (src/parser/scanner.ll, src/parser/parser.yy).  On failure
 * *out is NULL and the message is copied to err. */
mbg_status mbg_model_parse(const char* bnd_text, const char* cfg_text, mbg_model** out, char* err,
                           size_t errlen);
/* Same, reading the two files. */
mbg_status mbg_model_parse_files(const char* bnd_path, const char* cfg_path, mbg_model** out, char* err,
                                 size_t errlen);
void mbg_model_free(mbg_model* m);
/* A copy of a model (sharing its parsed form) with the same variant overrides; cheap, for
 * building many variants of one network. */
mbg_status mbg_model_clone(const mbg_model* m, mbg_model** out);

/* Model facts (driver::nodes / constants, src/parser/driver.h:41-43). */
int mbg_model_num_nodes(const mbg_model* m);
const char* mbg_model_node_name(const mbg_model* m, int i);
float mbg_model_node_istate(const mbg_model* m, int i);
int mbg_model_node_is_internal(const mbg_model* m, int i); /* node_t::is_internal, parse_types.cpp:20-22 */
int mbg_model_node_is_generic(const mbg_model* m, int i);  /* 0: "@logic ? U : 0" fast form */
int mbg_model_num_readouts(const mbg_model* m);            /* non-internal nodes */
/* A run constant (discrete_time, sample_count, max_time, time_tick, seed_pseudorandom) or
 * any other constant of the .cfg.  MBG_ERR_ARG if absent. */
mbg_status mbg_model_constant(const mbg_model* m, const char* name, double* out);

/* Per-node rate_up / rate_down multipliers and istate override for the next runs of this
 * model (variants: rate perturbations, knockouts = istate 0 and rate_up 0).  Any pointer may
 * be NULL to restore the model's own value.  No reference counterpart (README.md:33 lists
 * mutant analysis as future work). */
mbg_status mbg_model_set_variant(mbg_model* m, const float* mult_up, const float* mult_down,
                                 const float* istate);
/* The overrides in force: multipliers (1 where none is set) and initial-state probabilities
 * (the model's own where none is set), N values each; any pointer may be NULL. */
mbg_status mbg_model_get_variant(const mbg_model* m, float* mult_up, float* mult_down, float* istate);

/* The FP32 transition rate of every node at a state (state_words = ceil(N/64) words, node i
 * = bit i % 64 of word i / 64), from the tables the kernel reads (truth tables, logic
 * programs, generic rate tables), without variant multipliers: the host form of the
 * generated compute_transition_rates (src/generator.cpp:124-144). */
mbg_status mbg_model_transition_rates(const mbg_model* m, const uint64_t* state, float* rates);

/* The hipRTC source generated for this model (replaces generator::generate_code,
 * src/generator.cpp:11-55).  Writes at most cap bytes including the NUL; *len gets the full
 * length.  trace != 0 adds per-trajectory summaries. */
mbg_status mbg_model_generate_source(const mbg_model* m, int trace, char* buf, size_t cap, size_t* len);

/* ---------------------------------------------------------------- device context */

/* Creates a context on a HIP device (replaces kernel_compiler::kernel_compiler, which
 * creates a CUDA context on device 0, src/kernel_compiler.cpp:46-52). */
mbg_status mbg_create(int device, mbg_ctx** out);
void mbg_destroy(mbg_ctx* c);
const char* mbg_last_error(const mbg_ctx* c);

/* JIT-compiles (hipRTC, gfx950) and loads the trajectory kernel for this model's shape;
 * cached per shape.  Replaces kernel_compiler::compile_simulation
 * (src/kernel_compiler.cpp:61-164). */
mbg_status mbg_compile(mbg_ctx* c, const mbg_model* m, int trace);

typedef struct mbg_run_params {
    uint64_t n_trajectories; /* normalisation count n; 0 = sample_count */
    uint64_t first_id;       /* simulate trajectory ids [first_id, first_id + count) */
    uint64_t count;          /* 0 = n_trajectories */
    uint64_t seed;           /* Philox key */
    int use_model_seed;      /* 1: take seed from seed_pseudorandom */
    uint32_t variant_id;     /* Philox counter word 3 (distinguishes variants) */
    int trace;               /* 1: record per-trajectory summaries (count entries) */
    int block;               /* threads per block (multiple of 64); 0 = default */
    int grid;                /* blocks; 0 = fill the device */
    int compat_batch_resume; /* 1: reproduce the reference's 100-record launches (no flip after
                                every 99th step, a fresh waiting time for the same state next;
                                src/jit_kernels/simulation.cu:104-147); 0: uninterrupted Gillespie */
    int error_columns;       /* 1: also accumulate what the ErrorTH / ErrorProba columns need (the
                                reference writes 0 there, src/statistics/window_average_small.cpp:
                                211,235,245): per trajectory and window its time in each projected
                                state and its TH x time, squared and summed (exact integers) */
} mbg_run_params;

/* Simulates and accumulates the statistics (replaces simulation_runner::run_simulation,
 * src/simulation_runner.cu:54-180, with stats_composite::process_batch/finalize,
 * src/statistics/stats_composite.cpp:5-20).  Synchronous. */
mbg_status mbg_run(mbg_ctx* c, const mbg_model* m, const mbg_run_params* p);

/* Measured launch shape (no reference counterpart; the reference compiles one fixed launch
 * configuration): times candidate shapes on `sample` trajectories of the model (0: the
 * model's count, at most four per lane of a full chip, at least 262,144), by coordinate descent
 * over table placement, compact tables, register or LDS rate sums (or the lanes' whole state,
 * or their active sets, in HBM), block size, dependents per round and per iteration,
 * search width and multiply-add sum updates, and
 * keeps the fastest for later runs of models of the same structure (explicit blocks in
 * mbg_run_params override it).  Results never depend on the shape.  Discards the last run. */
typedef struct mbg_tune_info {
    int block, tables_in_lds, wsum_reg, compact, dep_batch, scan, wsmad, trials;
    uint64_t sample;
    double best_ms, default_ms;   /* kernel time of the chosen and of the untuned shape */
    int dep_cap;                  /* dependents re-evaluated per loop iteration (0: all) */
    int state_hbm;                /* 1: the lanes' state in HBM instead of LDS, 2: only the active sets */
} mbg_tune_info;
mbg_status mbg_autotune(mbg_ctx* c, const mbg_model* m, uint64_t sample, int trace, mbg_tune_info* out);

/* Shape of the last run's tables: W windows, K readouts (2^K projected states), fixed-point
 * exponents (cell = sum rint(t * 2^fp), th = sum rint(TH * t * 2^ft); discrete: cell is a
 * count), number of trajectories n, transitions taken, kernel time in ms. */
typedef struct mbg_run_info {
    int windows, readouts, prob_exponent, th_exponent, rate_exponent, state_words;
    int discrete;
    uint64_t n_trajectories, simulated, transitions, fixed_point_trajectories;
    double kernel_ms;          /* mbg_simulate time (HIP events), all batches */
    float time_tick, max_time;
    int block, tables_in_lds, lds_bytes, blocks_per_cu;  /* launch shape chosen for the run */
    int state_hbm;             /* lanes' state in HBM: 0 none (LDS), 1 all of it, 2 the active sets */
    int error_columns;         /* the run accumulated the error columns' sums of squares */
    int sq_exponent, th_sq_exponent; /* their units: 2^-sq_exponent of a squared window fraction
                                (count when discrete), 2^-th_sq_exponent of a squared window TH */
} mbg_run_info;
mbg_status mbg_get_run_info(const mbg_ctx* c, mbg_run_info* out);

/* Window tables (window_average_small_stats::finalize, src/statistics/window_average_small.cpp:
 * 109-140): prob[W * 2^K] and th[W] exact fixed-point sums; either pointer may be NULL. */
mbg_status mbg_get_probtraj_raw(const mbg_ctx* c, uint64_t* prob, uint64_t* th);
/* Normalised: P = cell / (n * tick) (or count / n), TH likewise. */
mbg_status mbg_get_probtraj(const mbg_ctx* c, double* prob, double* th);
/* Final projected states, counts[2^K] (final_states_stats, src/statistics/final_states.cpp:45-66). */
mbg_status mbg_get_final_states(const mbg_ctx* c, uint64_t* counts);
/* The non-zero window cells, window-major with projected states ascending: window index,
 * state (bit j = readout j) and exact fixed-point value (count when discrete), for any number
 * of readouts.  Runs with more than 20 readouts (the reference's limit, src/main.cpp:142-146)
 * keep their cells in a device hash table and have only this form: the dense getters above
 * return MBG_ERR_LIMIT for them.  *n gets the number of cells; at most cap are written (any
 * array may be NULL). */
mbg_status mbg_get_probtraj_cells(const mbg_ctx* c, uint32_t* windows, uint64_t* states, uint64_t* cells, size_t cap,
                                  size_t* n);
/* The non-zero final-state counts, states ascending (any number of readouts). */
mbg_status mbg_get_final_cells(const mbg_ctx* c, uint64_t* states, uint64_t* counts, size_t cap, size_t* n);
/* Distinct fixed points in the reference order (most significant word first,
 * src/state.h:43-49; fixed_states_stats, src/statistics/fixed_states.cu:84-187): states
 * [cap][state_words] and counts [cap]; *n gets the number of distinct fixed points. */
mbg_status mbg_get_fixed_points(const mbg_ctx* c, uint64_t* states, uint64_t* counts, size_t cap, size_t* n);
/* Per-trajectory summaries of a trace run: [count][4] = flips, FNV-1a hash of the flipped
 * node indices, final time (FP64 bits), status (1 finished, 2 fixed point). */
mbg_status mbg_get_summaries(const mbg_ctx* c, uint64_t* out, size_t count);

/* Error columns of the last run (run with error_columns = 1), as MaBoSS prints them beside each
 * probability: the standard error of the mean over trajectories of each trajectory's fraction of
 * window w spent in projected state s (its step count when discrete), err_prob[W * 2^K], and of
 * each trajectory's window-average TH, err_th[W]: sqrt((E[f^2] - E[f]^2) / (n - 1)), 0 when n < 2.
 * Replaces the literal zeros of src/statistics/window_average_small.cpp:211,235,245.  Either
 * pointer may be NULL; MBG_ERR_NOT_BUILT if the run did not accumulate them. */
mbg_status mbg_get_probtraj_errors(const mbg_ctx* c, double* err_prob, double* err_th);
/* The exact sums behind them: prob_sq[W * 2^K] and th_sq[W] (units in mbg_run_info). */
mbg_status mbg_get_probtraj_squares_raw(const mbg_ctx* c, uint64_t* prob_sq, uint64_t* th_sq);
/* Above 20 readouts: the sums of squares of the non-zero window cells, in the order of
 * mbg_get_probtraj_cells. */
mbg_status mbg_get_probtraj_cell_squares(const mbg_ctx* c, uint64_t* squares, size_t cap, size_t* n);

/* Readout marginals of the last run: out[W * K] = probability that readout j is ON, averaged
 * over window w (the window table summed over the projected states with bit j set). */
mbg_status mbg_get_marginals(const mbg_ctx* c, double* out);

/* ---------------------------------------------------------------- variant sweep */

/* One variant of a sweep: a model (its topology and, through mbg_model_set_variant, its rate
 * multipliers, knockouts and initial conditions) with its own trajectory count and seed. */
typedef struct mbg_variant {
    const mbg_model* model;
    uint64_t n_trajectories; /* trajectories to simulate (and normalisation count); 0 = sample_count */
    uint64_t seed;
    int use_model_seed;      /* 1: seed_pseudorandom of the model */
    uint32_t variant_id;     /* Philox counter word 3 */
} mbg_variant;

/* Runs every variant in one launch: the models are grouped by state-word class (1, 2, 4, ...
 * 64-bit words), one persistent kernel per class drains its variants' work queues, and all
 * pre-passes and class kernels are captured into one HIP graph (replaces running the
 * reference's CLI once per variant; README.md:33 lists mutant analysis as future work).
 * Synchronous.  The variants' tables stay on the device until the next run. */
mbg_status mbg_run_variants(mbg_ctx* c, const mbg_variant* variants, int count);
/* The same with flags: MBG_SWEEP_ERROR_COLUMNS also accumulates every variant's error-column
 * sums (as mbg_run_params.error_columns does for one run). */
#define MBG_SWEEP_ERROR_COLUMNS 1
mbg_status mbg_run_variants_ex(mbg_ctx* c, const mbg_variant* variants, int count, int flags);
/* Sweep autotuner (no reference counterpart: the reference has no sweeps).  For each word class
 * that these variants form, runs the class's variants alone under candidate layouts (tables
 * staged per block in LDS or read from HBM, active sets in HBM or LDS, dependents re-evaluated
 * per loop iteration), keeps the fastest per class, and applies it to later sweeps that form the same
 * class (mbg_sweep_class_info.tuned).  Results never depend on the layout.  flags as for
 * mbg_run_variants_ex.  Leaves the last tuning run as the context's last sweep. */
mbg_status mbg_autotune_variants(mbg_ctx* c, const mbg_variant* variants, int count, int flags);

typedef struct mbg_sweep_info {
    int variants, classes;
    uint64_t trajectories, transitions;
    double graph_ms;     /* the graph launch (HIP events), pre-passes and kernels included */
    double setup_ms;     /* host table building, uploads and graph construction */
} mbg_sweep_info;
mbg_status mbg_get_sweep_info(const mbg_ctx* c, mbg_sweep_info* out);
/* Word class i (0 <= i < classes) of the last sweep: its launch shape and what it ran.  No
 * reference counterpart (the reference runs one model per process). */
typedef struct mbg_sweep_class_info {
    int words, discrete, variants;
    uint64_t trajectories, transitions;
    double kernel_ms;    /* the class kernel's span inside the graph (first block start to last end) */
    int block, blocks_per_cu, state_hbm, wsum_reg, sumw, dep_batch, dep_cap, scan, th_lds;
    int tuned;           /* 1: the shape came from mbg_autotune_variants */
    int tables_lds;      /* 1: each block stages its current variant's tables in LDS */
} mbg_sweep_class_info;
mbg_status mbg_get_sweep_class_info(const mbg_ctx* c, int i, mbg_sweep_class_info* out);
/* Per-variant results of the last sweep (same layouts as the single-run getters). */
mbg_status mbg_get_variant_info(const mbg_ctx* c, int v, mbg_run_info* out);
mbg_status mbg_get_variant_probtraj_raw(const mbg_ctx* c, int v, uint64_t* prob, uint64_t* th);
/* Variant v's error-column sums of squares (a sweep run with MBG_SWEEP_ERROR_COLUMNS). */
mbg_status mbg_get_variant_probtraj_squares_raw(const mbg_ctx* c, int v, uint64_t* prob_sq, uint64_t* th_sq);
mbg_status mbg_get_variant_marginals(const mbg_ctx* c, int v, double* out);
mbg_status mbg_get_variant_final_states(const mbg_ctx* c, int v, uint64_t* counts);
mbg_status mbg_get_variant_fixed_points(const mbg_ctx* c, int v, uint64_t* states, uint64_t* counts, size_t cap,
                                        size_t* n);

/* Variants file: the overrides of a mutant / perturbation sweep, one block per variant.
 *
 *     variant NAME {                  // NAME: [A-Za-z_][A-Za-z0-9_-]*, unique
 *         NODE.knockout;              // istate 0, rate_up x 0
 *         NODE.overexpress;           // istate 1, rate_down x 0
 *         NODE.rate_up *= NUMBER;     // also rate_down; factors compose
 *         NODE.istate = NUMBER;       // in [0, 1]
 *         sample_count = INTEGER;     // default: the model's
 *         seed_pseudorandom = INTEGER;// default: the model's
 *     }
 *
 * For each block, out[i] gets a new model (a clone of base with the block's overrides applied
 * on top of base's own; the caller frees it with mbg_model_free), its trajectory count and
 * seed, and variant_id 0: every variant draws the same Philox numbers per trajectory id as a
 * plain run of its model (common random numbers), so an empty block reproduces mbg_run on base
 * bit for bit.  names (optional) gets each NUL-terminated name in 64-byte slots.  *count gets
 * the number of blocks; with out NULL the file is only checked.  MBG_ERR_PARSE with the line
 * in err on a syntax error, unknown node or bad value; MBG_ERR_LIMIT if more than cap blocks.
 * No reference counterpart (README.md:33 lists mutant analysis as future work). */
mbg_status mbg_variants_parse(const mbg_model* base, const char* text, size_t len, mbg_variant* out, char* names,
                              int cap, int* count, char* err, size_t errlen);
mbg_status mbg_variants_parse_file(const mbg_model* base, const char* path, mbg_variant* out, char* names, int cap,
                                   int* count, char* err, size_t errlen);
/* The reference's output files and stdout format for variant v of the last sweep (the same
 * writers as mbg_write_csv / mbg_write_text). */
mbg_status mbg_write_variant_csv(const mbg_ctx* c, int v, const char* prefix);
mbg_status mbg_write_variant_text(const mbg_ctx* c, int v, const char* path);

/* Output files of the reference: <prefix>_probtraj.csv (window_average_small_stats::write_csv,
 * src/statistics/window_average_small.cpp:181-250) and <prefix>_fp.csv
 * (fixed_states_stats::write_csv, src/statistics/fixed_states.cu:154-187). */
mbg_status mbg_write_csv(const mbg_ctx* c, const mbg_model* m, const char* prefix);
/* Stdout format (stats_composite::visualize: final_states.cpp:53-66, fixed_states.cu:141-152,
 * window_average_small.cpp:142-179) written to path ("-" = stdout). */
mbg_status mbg_write_text(const mbg_ctx* c, const mbg_model* m, const char* path);

#ifdef __cplusplus
}
#endif

#endif /* MBG_H */
