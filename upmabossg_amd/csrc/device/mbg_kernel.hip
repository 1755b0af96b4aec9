// Trajectory kernels + statistics stage for gfx950, compiled at run time by hipRTC.
//
// The code generator prepends, per model:
//   MBG_NW          64-bit state words per trajectory (ceil(N / 64))
//   MBG_BLOCK       threads per block of mbg_simulate (a multiple of 64)
//   MBG_DISCRETE    1 for discrete_time models
//   MBG_TRACE       1 to record per-trajectory summaries (flip count, flip hash)
//   MBG_GENERIC     1 if some node has a generic (tabulated) rate expression
//   MBG_TABLES_LDS  1 to stage the model tables in LDS (they fit), 0 to read them from HBM
//   MBG_MAX_IN      largest number of inputs of a node (gathers are unrolled up to 6)
//   MBG_DEP_BATCH   dependents evaluated per round of independent table reads
//   MBG_SCAN        candidate nodes per round of the in-word flip search
//   MBG_WSUM_REG    1: the word rate sums live in VGPRs (more lanes fit in LDS), 0: in LDS,
//                   2: in LDS, with sums of groups of MBG_SUMW words in VGPRs (wide models)
//   MBG_RATE32      1: every fixed-point rate fits 32 bits, the rate table is u32
//   MBG_MULTI       1: emit the variant-sweep kernel instead of the single-model one
//   MBG_SUMW        64-bit words per register rate-sum group (divides MBG_NW when
//                   MBG_WSUM_REG is 1; the last group may be short when it is 2)
//   MBG_COMPACT     1: 8-byte node records and 16-bit dependents (small models, no generic rates)
//   MBG_WSMAD       1: every rate is below 2^31 (register sum updates by multiply-add)
//   MBG_SPARSE      1: more than 20 readouts, window cells and final states in a hash table
//   MBG_PROG        1: some fast node reads more than 20 nodes (logic program, no truth table)
//   MBG_TH_LDS      1: partial sums of the window entropies in LDS (per block in single runs,
//                   per wave and variant in sweeps)
//   MBG_CELLS_LDS   1: also of the whole window table and final states (small tables)
//   MBG_ERRORS      1: also the sums of squares behind the ErrorProba / ErrorTH columns (each
//                   trajectory's time per projected state and TH x time over each window,
//                   squared); each lane keeps its current window's (state, time) list in HBM
//   MBG_GENCODE     1: the dependents' logic from code generated per node (appended to this
//                   source) instead of the truth tables (a comparison; the tables are faster)
//   MBG_SWEEP_XCD   sweep waves start at variant 0, the costliest per trajectory (2), at variants
//                   grouped by block and XCD (1), or spread per wave (0)
//   MBG_STATE_HBM   1: the lanes' x, act and word sums in an HBM buffer (lane-interleaved over
//                   the whole grid) instead of LDS, for models whose per-lane state would leave
//                   a CU with one or two waves (single runs only); 2: only act in that buffer,
//                   also in sweeps (a class's grid shares one buffer, lane by lane),
//                   x in LDS and the word sums in VGPRs (half the LDS per lane, so twice the
//                   resident waves; act is read once per step and once per rate change);
//                   3: act and the word sums in that buffer, x in LDS and the group sums in
//                   VGPRs (MBG_WSUM_REG 2; wide models, single runs: the input gathers of the
//                   dependents and the flips stay in LDS)
// and mbg_math.h.
//
// One trajectory per lane (the reference runs one CUDA thread per trajectory,
// src/jit_kernels/simulation.cu:77-164).  Unlike the reference's O(N) rate recompute
// and linear flip scan per step (src/generator.cpp:124-144, simulation.cu:9-21), each
// lane keeps
//   x[w]    the Boolean state,
//   act[w]  the active set (nodes with a non-zero rate),
//   wsum[w] the exact fixed-point rate sum of each 64-node word,
// x and act in LDS, lane-interleaved (element w of lane l at w*BLOCK + l: no two lanes of
// a wave share a bank whatever word each touches), wsum in VGPRs (MBG_WSUM_REG, read and
// written through unrolled selects) or in LDS the same way; for wide models all of it in an
// HBM buffer interleaved the same way over the grid (MBG_STATE_HBM).  A transition
// re-evaluates only the flipped node and its dependents from per-node truth tables and
// applies exact integer deltas, so the two-level search (word sums, then the active bits of
// one word) selects exactly the node the reference's linear scan of prefix sums would.  The
// dependents are re-evaluated at the start of the next loop iteration, at most MBG_DEP_CAP of
// them per iteration, so one lane with many dependents does not hold its whole wave.
//
// Initial states are built by a throughput-bound pre-pass (mbg_init_rates: the initial
// draws and every node's rate for 64 trajectories per block, staged through LDS, then the
// readout sums; above 128 words mbg_init_state draws the states first and mbg_init_readouts
// sums the readouts after) into an HBM staging area, so a lane that starts a trajectory only
// copies 3*NW + 4 words instead of evaluating all N nodes while its wave waits.
//
// Statistics (src/jit_kernels/window_average_small.cu:50-136, final_states.cu:9-21,
// src/statistics/fixed_states.cu:84-125) are accumulated per lane in registers for the
// current (window, projected state) cell and flushed with 64-bit integer atomics when the
// cell changes: the tables are exact sums, independent of scheduling.

typedef unsigned int u32;
typedef unsigned short u16;
typedef unsigned long long u64;
// four u32 read as one 16-byte access
#ifndef MBG_U32X4
typedef u32 mbg_u32x4 __attribute__((ext_vector_type(4)));
#endif

// Pointers into memory that no kernel writes while it runs (the model tables).
#ifndef CONST_AS
#define CONST_AS __attribute__((address_space(4)))
#endif
// A value the compiler must treat as unknown from here on (an empty asm that "writes" it), in
// a vector register (V) or a scalar one (S): keeps a computation in the form it is written in.
#ifndef MBG_OPAQUE_V
#define MBG_OPAQUE_V(x) asm volatile("" : "+v"(x))
#define MBG_OPAQUE_S(x) asm volatile("" : "+s"(x))
#endif
// The kernel's dynamic LDS, sized at launch.
#ifndef MBG_DYN_LDS
#define MBG_DYN_LDS(name) extern __shared__ u64 name[]
#endif

struct mbg_node {
    u32 in_off;               // first input in the inputs table
    u32 dep_off;              // first dependent in the deps table
    u32 tab_off;              // fast: bit offset into lut; generic: index into rtab
    unsigned short dep_cnt;   // dependents (nodes whose rate reads this node, itself included)
    unsigned char in_cnt;     // inputs other than the node itself
    unsigned char flags;      // bit 7: generic rate table; bit 6: listed among its own
                              // dependents; bits 0-5: readout index + 1 (0: internal)
};

// Words per staged initial state (the mbg_desc::stage rows).
#define STAGE_ROW (3 * MBG_NW + 4)

#define F_GEN 0x80u
#define F_SELF 0x40u
#define F_RO 0x3fu

#ifndef MBG_COMPACT
#define MBG_COMPACT 0
#endif
#ifndef MBG_SPARSE
#define MBG_SPARSE 0
#endif
#ifndef MBG_PROG
#define MBG_PROG 0
#endif
#ifndef MBG_TH_LDS
#define MBG_TH_LDS 0
#endif
#ifndef MBG_CELLS_LDS
#define MBG_CELLS_LDS 0
#endif
#ifndef MBG_ERRORS
#define MBG_ERRORS 0
#endif
#if !MBG_TH_LDS || MBG_SPARSE || MBG_MULTI || MBG_ERRORS  // the cell partials follow the entropy partials;
                                           // dense tables of single runs only
#undef MBG_CELLS_LDS
#define MBG_CELLS_LDS 0
#endif
#define PROG_IN 255u  // in_cnt of a node whose logic is a program
// Projected (readout) state: up to 20 bits in a dense table, up to 63 with MBG_SPARSE.
#if MBG_SPARSE
typedef unsigned long long sidx;
#else
typedef unsigned int sidx;
#endif
#ifndef MBG_WSMAD
#define MBG_WSMAD 0
#endif
#ifndef MBG_DEP_CAP
#define MBG_DEP_CAP 0
#endif
#ifndef MBG_STATE_HBM
#define MBG_STATE_HBM 0
#endif
#ifndef MBG_SWEEP_XCD
#define MBG_SWEEP_XCD 0
#endif
#ifndef MBG_GENCODE
#define MBG_GENCODE 0
#endif
// The active sets alone in HBM keep x alone in LDS: the rate sums must be in registers.
#if MBG_STATE_HBM == 2 && MBG_WSUM_REG != 1
#error "the active sets alone in HBM need register rate sums"
#endif
#if MBG_STATE_HBM == 3 && (MBG_WSUM_REG != 2 || MBG_MULTI)
#error "the active sets and word sums in HBM go with register group sums, in single runs"
#endif

// A node record decoded into registers (either table layout).
struct mbg_nd {
    u32 in_off, dep_off, tab_off, dep_cnt, in_cnt, flags;
};

// MBG_COMPACT layout: one u64 per node, bits 0-13 in_off, 14-27 dep_off, 28-43 tab_off,
// 44-51 dep_cnt, 52-55 in_cnt, 56 F_SELF, 57-61 readout index + 1 (no generic nodes); the
// dependents table holds u16 node (12 bits) | key bit position (4 bits).
__device__ __forceinline__ mbg_nd decode_compact(u64 v) {
    mbg_nd n;
    n.in_off = (u32)v & 0x3fffu;
    n.dep_off = (u32)(v >> 14) & 0x3fffu;
    n.tab_off = (u32)(v >> 28) & 0xffffu;
    n.dep_cnt = (u32)(v >> 44) & 0xffu;
    n.in_cnt = (u32)(v >> 52) & 0xfu;
    n.flags = (((u32)(v >> 56) & 1u) << 6) | ((u32)(v >> 57) & 0x1fu);
    return n;
}

__device__ __forceinline__ mbg_nd decode_wide(const mbg_node& r) {
    mbg_nd n;
    n.in_off = r.in_off;
    n.dep_off = r.dep_off;
    n.tab_off = r.tab_off;
    n.dep_cnt = r.dep_cnt;
    n.in_cnt = r.in_cnt;
    n.flags = r.flags;
    return n;
}

// Byte offsets of the model tables inside one blob (global memory, or LDS when staged).
struct mbg_layout {
    u32 nodes;     // mbg_node[N]
    u32 rates;     // u64[2N] (u32[2N] with MBG_RATE32): fixed-point [rate_up, rate_down]
    u32 deps;      // u32[]: node id | (key bit position << 24)
    u32 inputs;    // u16[]: node ids
    u32 lut;       // u32[]: logic truth tables, key = x_i | inputs << 1
    u32 readouts;  // u32[K]: non-internal node ids in node order
    u32 terms;     // u64[2K]: entropy terms rint(r log2(r) 2^sigma) of each readout's U and D
    u32 gen_mask;  // u64[NW]: nodes with a generic rate table
    u32 bytes;     // blob size
    u32 prog;      // u32[]: logic programs of nodes with in_cnt 255 ([length, ops...])
    u32 onehot;    // u32[16][16] identity: row g holds the factors [i == g] of a rate-sum update
};

struct mbg_desc {
    const char* blob;         // model tables (layout above)
    const u64* rtab;          // fixed-point rate per key of generic nodes
    const float* istate;      // initial ON probability per node
    u64* stage;               // [batch][STAGE_ROW]: x, act, wsum (NW words each) of each trajectory's
                              // initial state, then its projected state, readout rate sum,
                              // readout entropy-term sum and active-readout count
    u64* prob;                // [W][2^K] fixed-point sojourn time (counts when discrete)
    u64* th;                  // [W] fixed-point transition entropy * time
    u64* fin;                 // [2^K] final projected states
    u64* fp_states;           // [fp_cap][NW] fixed-point states, append order
    u64* fp_count;
    u64* summ;                // [summ_cap][4] flips, flip hash, final time bits, status
    u64* work_counter;        // next trajectory of the batch to start
    u64* flip_total;          // transitions taken
    u64 fp_cap;
    u64 summ_cap;
    u64 nsteps;               // discrete: steps per trajectory
    u64 batch;                // trajectories in this batch
    u64 batch_first;          // id of the batch's first trajectory
    u64 first_id;             // id of the run's first trajectory (summary index base)
    u64 seed;
    double maxt, tick, inv_tick, rscale;
    mbg_layout lay;
    u32 variant_id;
    int N, K, W, FP, FT, sigma;
    int compat;               // 1: no flip after every 99th step (the reference's 100-record launches)
    int pad_;
    // MBG_SPARSE (more than 20 readouts): window cells and final states in one open-addressing
    // table instead of prob / fin; key = window << K | projected state (window W: final states)
    u64* cells;               // [cell_mask + 1] keys, ~0 = empty
    u64* cell_vals;           // [cell_mask + 1] fixed-point sums / counts
    u64* cell_miss;           // inserts that found no free slot within the probe limit
    u64 cell_mask;
    // MBG_STATE_HBM: the lanes' state, element (array, word) of lane t at
    // (array * NW + word) * lane_stride + t, for t below lane_stride (at least the grid's threads)
    u64* lane_state;
    u32 lane_stride;
    u32 pad2_;
    // MBG_ERRORS: sums over trajectories of squared per-trajectory window totals, prob2 [W][2^K]
    // (cell_vals2 beside cell_vals when sparse) and th2 [W], in units 2^-FQ (2^-FQT) of the
    // squared window fraction (window-average TH), a total v scaling to it by sq_scale
    // (sq_scale_th); each lane's list of the current window's (state, time) entries, entry j of
    // grid thread t at j * err_stride + t, at most err_cap entries (err_miss counts overflows)
    u64* prob2;
    u64* th2;
    u64* cell_vals2;
    u64* err_s;
    u64* err_v;
    u64* err_miss;
    double sq_scale, sq_scale_th;
    int FQ, FQT;
    u32 err_cap, err_stride;
};

#if MBG_COMPACT
#define TNODE(i) decode_compact(((const u64*)(tb + D.lay.nodes))[i])
#else
#define TNODE(i) decode_wide(((const mbg_node*)(tb + D.lay.nodes))[i])
#endif
#if MBG_RATE32
#define TRATE(i) ((u64)((const u32*)(tb + D.lay.rates))[i])
#else
#define TRATE(i) (((const u64*)(tb + D.lay.rates))[i])
#endif
#if MBG_COMPACT
#define TDEP(i) dep_entry(((const unsigned short*)(tb + D.lay.deps))[i])
__device__ __forceinline__ u32 dep_entry(u32 v) { return (v & 0xfffu) | ((v >> 12) << 24); }
#else
#define TDEP(i) (((const u32*)(tb + D.lay.deps))[i])
#endif
#define TIN(i) ((u32)((const u16*)(tb + D.lay.inputs))[i])
#define TPROG(i) (((const u32*)(tb + D.lay.prog))[i])

#if MBG_PROG
// Logic of a node that reads more than 20 nodes: a postfix program over a stack of bits held
// in one u64 (op format in csrc/model.cpp: push node bit / constant / clause, AND, OR, XOR,
// NOT, SELECT), word(w) giving state word w.  Every op runs the same branch-free
// instructions, so lanes evaluating different nodes only differ in their loop count.  Each op
// reads one state word: the clause's, or the pushed node's (for other ops the index is node
// 0, 1 or 2, always valid).  The 5 words after an op are read unconditionally (the table is
// padded past its last op).
template <class Word>
__device__ __forceinline__ u32 eval_prog(const mbg_desc& D, const char* tb, u32 off, Word word) {
    const u32 end = off + 1u + TPROG(off);
    u64 st = 0;
    for (u32 p = off + 1u; p < end;) {
        const u32 op = TPROG(p);
        const u32 kind = op >> 30;
        const bool clause = kind == 3u && ((op >> 29) & 1u);
        const u32 cw = TPROG(p + 1u);
        const u64 cm = (u64)TPROG(p + 2u) | ((u64)TPROG(p + 3u) << 32);
        const u64 cv = (u64)TPROG(p + 4u) | ((u64)TPROG(p + 5u) << 32);
        const u32 j = op & 0xffffu;
        const u64 wv = word(clause ? cw : (j >> 6));
        const u32 a = (u32)st & 1u, b = (u32)(st >> 1) & 1u, c = (u32)(st >> 2) & 1u;
        const u32 lit = clause ? ((u32)((wv & cm) == cv) ^ ((op >> 28) & 1u))
                               : (kind == 0u ? (((u32)(wv >> (j & 63)) & 1u) ^ ((op >> 16) & 1u)) : (op & 1u));
        const u32 f = op & 3u;
        const u32 bin = f == 0u ? (a & b) : (f == 1u ? (a | b) : (a ^ b));
        const u64 pushv = (st << 1) | (u64)lit;
        const u64 binv = ((st >> 2) << 1) | (u64)bin;
        const u64 selv = ((st >> 3) << 1) | (u64)(c ? b : a);
        const u64 notv = st ^ 1ull;
        st = (kind == 0u || kind == 3u) ? pushv : (kind == 1u ? binv : ((op & 1u) ? selv : notv));
        p += clause ? 6u : 1u;
    }
    return (u32)st & 1u;
}
#define IN_CNT(n) ((n).in_cnt == PROG_IN ? 0u : (n).in_cnt)
#else
#define IN_CNT(n) ((n).in_cnt)
#endif
#define TLUT(i) (((const u32*)(tb + D.lay.lut))[i])
#define TRO(i) (((const u32*)(tb + D.lay.readouts))[i])
#define TTERM(i) (((const u64*)(tb + D.lay.terms))[i])
#define TGEN(i) (((const u64*)(tb + D.lay.gen_mask))[i])

// Lane state: element w of a lane's array at w * LSTRIDE + tid (no two lanes of a wave share
// an LDS bank whatever word each touches; in HBM, consecutive lanes share cache lines).
#if MBG_STATE_HBM == 1
#define LSTRIDE D.lane_stride
#else
#define LSTRIDE MBG_BLOCK
#endif
#define LANE_X(w) lane[(0 * MBG_NW + (w)) * LSTRIDE + tid]
#if MBG_STATE_HBM >= 2
// act in HBM, word w of grid thread atid at w * lane_stride + atid
#define LANE_A(w) D.lane_state[(u64)(w) * D.lane_stride + atid]
#else
#define LANE_A(w) lane[(1 * MBG_NW + (w)) * LSTRIDE + tid]
#endif
#if MBG_WSUM_REG
// rate sums of groups of MBG_SUMW words in registers: every access is an unrolled select
// over the groups (at most 16)
#if MBG_WSUM_REG == 2
#define NSUM ((MBG_NW + MBG_SUMW - 1) / MBG_SUMW)
#else
#define NSUM (MBG_NW / MBG_SUMW)
#endif
#define LANE_S(g) ws[g]
#if MBG_WSMAD && MBG_TABLES_LDS
// every rate change fits an i32: ws[i] += [i == g] * d as one 32 x 32 + 64 multiply-add, the
// factors [i == g] read as row g of the identity table staged in LDS with the model (four
// 16-byte reads on the LDS pipe instead of one bit-field extract per sum on the VALU)
#define WS_GADD(g, d)                                                      \
    do {                                                                   \
        const mbg_u32x4* row_ = (const mbg_u32x4*)(tb + D.lay.onehot) + (g) * 4u; \
        const long long d_ = (long long)(int)(u32)(d);                     \
        _Pragma("unroll") for (int q_ = 0; q_ < (NSUM + 3) / 4; q_++) {    \
            const mbg_u32x4 f4_ = row_[q_];                                \
            const u32 f_[4] = {f4_[0], f4_[1], f4_[2], f4_[3]};            \
            _Pragma("unroll") for (int e_ = 0; e_ < 4; e_++) {             \
                const int i_ = q_ * 4 + e_;                                \
                if (i_ < NSUM)                                             \
                    ws[i_] = (u64)((long long)ws[i_] + (long long)(int)f_[e_] * d_); \
            }                                                              \
        }                                                                  \
    } while (0)
#elif MBG_WSMAD
// every rate change fits an i32: ws[i] += [i == g] * d as one 32 x 32 + 64 multiply-add, the
// factor [i == g] as one bit-field extract of the one-hot word (hidden from the compiler, which
// would otherwise rebuild each factor as a compare and a select)
#define WS_GADD(g, d)                                                      \
    do {                                                                   \
        u32 oh_ = 1u << (g);                                               \
        MBG_OPAQUE_V(oh_);                                                 \
        const long long d_ = (long long)(int)(u32)(d);                     \
        _Pragma("unroll") for (int i_ = 0; i_ < NSUM; i_++) {              \
            u32 f_ = (oh_ >> i_) & 1u;                                     \
            MBG_OPAQUE_V(f_); /* keep the multiply-add */                  \
            ws[i_] = (u64)((long long)ws[i_] + (long long)(int)f_ * d_);   \
        }                                                                  \
    } while (0)
#else
#define WS_GADD(g, d)                                                      \
    do {                                                                   \
        const u32 g_ = (g);                                                \
        const u64 d_ = (d);                                                \
        _Pragma("unroll") for (int i_ = 0; i_ < NSUM; i_++) ws[i_] += g_ == (u32)i_ ? d_ : 0ull; \
    } while (0)
#endif
#if MBG_WSUM_REG == 2
#if MBG_STATE_HBM == 3
// the word sums themselves in HBM after act, word w of grid thread atid at (NW + w) * lane_stride + atid
#define LANE_W(w) D.lane_state[(u64)(MBG_NW + (w)) * D.lane_stride + atid]
#else
// the word sums themselves in LDS (lane-interleaved like x and act)
#define LANE_W(w) lane[(2 * MBG_NW + (w)) * LSTRIDE + tid]
#endif
#define WS_ADD(w, d)                                                       \
    do {                                                                   \
        const u32 w2_ = (w);                                               \
        const u64 dd_ = (d);                                               \
        LANE_W(w2_) = LANE_W(w2_) + dd_;                                   \
        WS_GADD(w2_ / MBG_SUMW, dd_);                                      \
    } while (0)
#else
#define WS_ADD(w, d) WS_GADD((w) / MBG_SUMW, d)
#endif
#else
#undef MBG_SUMW
#define MBG_SUMW 1
#define NSUM MBG_NW
#define LANE_S(w) lane[(2 * MBG_NW + (w)) * LSTRIDE + tid]
#define WS_ADD(w, d) LANE_S(w) = LANE_S(w) + (d)
#endif

// ------------------------------------------------------------------ pre-pass kernels

// Initial state word w of trajectory id: node i is ON iff the i-th uniform of the
// trajectory's stream 0 is <= istate_i (simulation.cu:34-71).
__device__ __forceinline__ u64 initial_word(const mbg_desc& D, u32 id, int w) {
    u64 word = 0;
    for (int q = 0; q < 16; q++) {
        const int i0 = w * 64 + q * 4;
        if (i0 >= D.N) break;
        u32 o[4];
        mbg_philox((u32)(i0 >> 2), 0u, id, D.variant_id, (u32)D.seed, (u32)(D.seed >> 32), o);
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int i = i0 + r;
            if (i < D.N && mbg_unit_f32(o[r]) <= D.istate[i]) word |= 1ull << (i & 63);
        }
    }
    return word;
}

// Pass 1 (models above 128 words), one thread per (trajectory, word): initial state bits.
extern "C" __global__ __launch_bounds__(256) void mbg_init_state(mbg_desc D) {
    const u64 gid = (u64)blockIdx.x * 256u + threadIdx.x;
    if (gid >= D.batch * MBG_NW) return;
    const u64 t = gid / MBG_NW;
    const int w = (int)(gid - t * MBG_NW);
    D.stage[t * STAGE_ROW + w] = initial_word(D, (u32)(D.batch_first + t), w);
}

__device__ __forceinline__ u64 rate_from_key(const mbg_desc& D, const char* tb, const mbg_nd& n, u32 i, u32 key) {
#if MBG_GENERIC
    if (n.flags & 0x80) return D.rtab[n.tab_off + key];
#endif
    const u32 bp = n.tab_off + key;
    const u32 logic = (TLUT(bp >> 5) >> (bp & 31)) & 1u;
    const u32 xi = key & 1u;
    return xi == logic ? 0ull : TRATE(2 * i + xi);
}

// The same for any node, with word(w) reading the state for a logic program.
template <class Word>
__device__ __forceinline__ u64 rate_of(const mbg_desc& D, const char* tb, const mbg_nd& n, u32 i, u32 key, Word word) {
#if MBG_PROG
    if (n.in_cnt == PROG_IN) {
        const u32 xi = key & 1u;
        return xi == eval_prog(D, tb, n.tab_off, word) ? 0ull : TRATE(2 * i + xi);
    }
#endif
    (void)word;
    return rate_from_key(D, tb, n, i, key);
}

// Entropy term rint(r log2(r) 2^sigma) of readout ro (index) at rate r produced with own state
// bit xi: a table entry for fast nodes, computed for generic ones.
__device__ __forceinline__ u64 entropy_term(const mbg_desc& D, const char* tb, const mbg_nd& n, u32 ro, u32 xi,
                                            u64 r) {
    if (!r) return 0ull;
#if MBG_GENERIC
    if (n.flags & F_GEN) {
        const double rd = (double)r;
        return (u64)rint(ldexp(rd * mbg_log2(rd), D.sigma));
    }
#else
    (void)D;
    (void)n;
#endif
    return TTERM(2 * ro + xi);
}

// Readout part of a state (word(w) gives state word w): projected state, readout rate sum,
// entropy-term sum and active-readout count, over readouts j0, j0 + step, ...
template <class Word>
__device__ __forceinline__ void readout_sums(const mbg_desc& D, const char* tb, int j0, int step, Word word, u64& s,
                                             u64& R, u64& S, u64& act) {
    s = R = S = act = 0;
    for (int j = j0; j < D.K; j += step) {
        const u32 n = TRO(j);
        const mbg_nd nd = TNODE(n);
        u32 key = (u32)(word(n >> 6) >> (n & 63)) & 1u;
        for (u32 m = 0; m < IN_CNT(nd); m++) {
            const u32 q = TIN(nd.in_off + m);
            key |= ((u32)(word(q >> 6) >> (q & 63)) & 1u) << (m + 1);
        }
        const u64 r = rate_of(D, tb, nd, n, key, word);
        const u32 xi = key & 1u;
        s |= (u64)xi << j;
        R += r;
        S += entropy_term(D, tb, nd, (u32)j, xi, r);
        act += r != 0 ? 1ull : 0ull;
    }
}

// Initial states, active sets and word rate sums.  A block takes 64 trajectories: it draws
// their state words into LDS (lane-interleaved) and the staging area, then each wave sweeps
// whole words for the 64 trajectories, one per lane, so the table reads of a word's 64
// nodes are uniform across the wave and the state reads never conflict.  Models above 128
// words run mbg_init_state first and use the per-(trajectory, word) variant below.
#define INIT_TRAJ 64
#if MBG_NW <= 128
__device__ __forceinline__ void init_rates_block(const mbg_desc& D, u64 t0) {
    __shared__ u64 xs[MBG_NW * INIT_TRAJ];
    const u32 tid = threadIdx.x;
    const u32 lane = tid & 63u;
    for (u32 idx = tid; idx < MBG_NW * INIT_TRAJ; idx += 256) {
        const u32 w = idx / INIT_TRAJ, l = idx % INIT_TRAJ;
        u64 word = 0;
        if (t0 + l < D.batch) {
            word = initial_word(D, (u32)(D.batch_first + t0 + l), (int)w);
            D.stage[(t0 + l) * STAGE_ROW + w] = word;
        }
        xs[idx] = word;
    }
    __syncthreads();
    const u64 t = t0 + lane;
    const char* tb = D.blob;
    // Eight nodes per round: their keys are gathered first and their table reads issued
    // together, so the wave waits for one round of loads instead of one per node.  Node
    // indices are wave-uniform (scalar table reads); past the last node they are clamped to
    // it and the result discarded.
    for (int w = (int)__builtin_amdgcn_readfirstlane(tid >> 6); w < MBG_NW; w += 4) {
        u64 a = 0, sum = 0;
        const int nb = min(64, D.N - w * 64);
        for (int b0 = 0; b0 < nb; b0 += 8) {
            u32 key[8], tab[8], inc[8];
            u64 rup[8], rdn[8];
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const u32 i = (u32)min(w * 64 + b0 + q, D.N - 1);
                // node record and input list through the constant address space: the
                // index is wave-uniform, so these become scalar loads
#if MBG_COMPACT
                const mbg_nd nd = decode_compact(((const CONST_AS u64*)(tb + D.lay.nodes))[i]);
                const u32 in_off = nd.in_off, in_cnt = nd.in_cnt;
                tab[q] = nd.tab_off;
                inc[q] = in_cnt;
#else
                const CONST_AS u32* nr = (const CONST_AS u32*)(tb + D.lay.nodes) + 4 * i;
                const u32 in_off = nr[0], in_raw = nr[3] >> 16 & 0xffu;
#if MBG_PROG
                const u32 in_cnt = in_raw == PROG_IN ? 0u : in_raw;
#else
                const u32 in_cnt = in_raw;
#endif
                tab[q] = nr[2];
                inc[q] = in_raw;
#endif
#if !MBG_GENERIC
#if MBG_RATE32
                const CONST_AS u32* rt = (const CONST_AS u32*)(tb + D.lay.rates);
#else
                const CONST_AS u64* rt = (const CONST_AS u64*)(tb + D.lay.rates);
#endif
                rup[q] = rt[2 * i];
                rdn[q] = rt[2 * i + 1];
#endif
                u32 k = (u32)(xs[(i >> 6) * INIT_TRAJ + lane] >> (i & 63)) & 1u;
                const CONST_AS u32* ins = (const CONST_AS u32*)(tb + D.lay.inputs);
#if MBG_MAX_IN <= 6
                // the node's (up to 6) 16-bit input ids in one 16-byte scalar load; reads
                // past its inputs land in the next node's entries or the table's padding
                // and are masked off
                const u32 b4 = in_off >> 1, odd = in_off & 1u;
                const u32 iw[4] = {ins[b4], ins[b4 + 1], ins[b4 + 2], ins[b4 + 3]};
#pragma unroll
                for (u32 m = 0; m < MBG_MAX_IN; m++) {
                    const u32 loc = odd + m;
                    const u32 j = (iw[loc >> 1] >> ((loc & 1u) * 16u)) & 0xffffu;
                    const u32 b = (u32)(xs[(j >> 6) * INIT_TRAJ + lane] >> (j & 63)) & 1u;
                    k |= (m < in_cnt ? b : 0u) << (m + 1);
                }
#else
                for (u32 m = 0; m < in_cnt; m++) {
                    const u32 e = in_off + m;
                    const u32 j = (ins[e >> 1] >> ((e & 1u) * 16u)) & 0xffffu;
                    k |= ((u32)(xs[(j >> 6) * INIT_TRAJ + lane] >> (j & 63)) & 1u) << (m + 1);
                }
#endif
                key[q] = k;
            }
            u64 r[8];
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const u32 i = (u32)min(w * 64 + b0 + q, D.N - 1);
                auto xsword = [&](u32 w) { return xs[w * INIT_TRAJ + lane]; };
                (void)xsword;
#if MBG_GENERIC
                u64 v = rate_of(D, tb, TNODE(i), i, key[q], xsword);
#else
                // fast form: the truth-table bit per lane, the two rates of the (uniform) node
                u32 logic;
#if MBG_PROG
                if (inc[q] == PROG_IN) logic = eval_prog(D, tb, tab[q], xsword);
                else
#endif
                {
                    const u32 bp = tab[q] + key[q];
                    logic = (TLUT(bp >> 5) >> (bp & 31)) & 1u;
                }
                const u32 xi = key[q] & 1u;
                const u64 v = (xi ? rdn[q] : rup[q]) & (0ull - (u64)(xi != logic));
                (void)i;
#endif
                r[q] = b0 + q < nb ? v : 0ull;
            }
#pragma unroll
            for (int q = 0; q < 8; q++) {
                if (r[q]) {
                    a |= 1ull << (b0 + q);
                    sum += r[q];
                }
            }
        }
        if (t < D.batch) {
            D.stage[t * STAGE_ROW + MBG_NW + w] = a;
            D.stage[t * STAGE_ROW + 2 * MBG_NW + w] = sum;
        }
    }
    // the readout part of each staged state: wave q sums readouts q, q + 4, ... of the
    // trajectory in its lane, then the first wave adds the four partial sums
    __shared__ u64 rp[4][4][INIT_TRAJ];
    {
        const u32 q = tid >> 6;
        u64 s, R, S, act;
        readout_sums(D, tb, (int)q, 4, [&](u32 w) { return xs[w * INIT_TRAJ + lane]; }, s, R, S, act);
        rp[q][0][lane] = s;
        rp[q][1][lane] = R;
        rp[q][2][lane] = S;
        rp[q][3][lane] = act;
    }
    __syncthreads();
    if (tid < INIT_TRAJ && t0 + tid < D.batch) {
        u64* row = D.stage + (t0 + tid) * STAGE_ROW + 3 * MBG_NW;
        row[0] = rp[0][0][tid] | rp[1][0][tid] | rp[2][0][tid] | rp[3][0][tid];
#pragma unroll
        for (int f = 1; f < 4; f++) row[f] = rp[0][f][tid] + rp[1][f][tid] + rp[2][f][tid] + rp[3][f][tid];
    }
}

extern "C" __global__ __launch_bounds__(256) void mbg_init_rates(mbg_desc D) {
    init_rates_block(D, (u64)blockIdx.x * INIT_TRAJ);
}

#if MBG_MULTI
// The same pass for every variant of a sweep class in one launch: block (x, v) takes
// trajectories 64x .. 64x + 63 of variant v.
extern "C" __global__ __launch_bounds__(256) void mbg_init_rates_multi(const CONST_AS mbg_desc* descs) {
    const mbg_desc& D = *(const mbg_desc*)(descs + blockIdx.y);
    const u64 t0 = (u64)blockIdx.x * INIT_TRAJ;
    if (t0 >= D.batch) return;  // the whole block: it holds no barrier yet
    init_rates_block(D, t0);
}
#endif
#else
extern "C" __global__ __launch_bounds__(256) void mbg_init_rates(mbg_desc D) {
    const u64 gid = (u64)blockIdx.x * 256u + threadIdx.x;
    if (gid >= D.batch * MBG_NW) return;
    const u64 t = gid / MBG_NW;
    const int w = (int)(gid - t * MBG_NW);
    const u64* x = D.stage + t * STAGE_ROW;
    const char* tb = D.blob;
    u64 a = 0, sum = 0;
    for (int b = 0; b < 64; b++) {
        const u32 i = (u32)(w * 64 + b);
        if ((int)i >= D.N) break;
        const mbg_nd n = TNODE(i);
        u32 key = (u32)(x[i >> 6] >> (i & 63)) & 1u;
        for (u32 m = 0; m < IN_CNT(n); m++) {
            const u32 j = TIN(n.in_off + m);
            key |= ((u32)(x[j >> 6] >> (j & 63)) & 1u) << (m + 1);
        }
        const u64 r = rate_of(D, tb, n, i, key, [&](u32 w) { return x[w]; });
        if (r) {
            a |= 1ull << b;
            sum += r;
        }
    }
    D.stage[t * STAGE_ROW + MBG_NW + w] = a;
    D.stage[t * STAGE_ROW + 2 * MBG_NW + w] = sum;
}
#endif

// Models above 128 words, pass 3, one thread per staged trajectory: the readout part of its
// initial state at the end of its stage row (smaller models get it from mbg_init_rates), so that
// a lane starting the trajectory reads four words instead of looping over the readouts while the
// rest of its wave waits.
__device__ __forceinline__ void init_readouts_one(const mbg_desc& D, u64 t) {
    u64* row = D.stage + t * STAGE_ROW;
    u64 s, R, S, act;
    readout_sums(D, D.blob, 0, 1, [&](u32 w) { return row[w]; }, s, R, S, act);
    row[3 * MBG_NW + 0] = s;
    row[3 * MBG_NW + 1] = R;
    row[3 * MBG_NW + 2] = S;
    row[3 * MBG_NW + 3] = act;
}

extern "C" __global__ __launch_bounds__(256) void mbg_init_readouts(mbg_desc D) {
    const u64 t = (u64)blockIdx.x * 256u + threadIdx.x;
    if (t < D.batch) init_readouts_one(D, t);
}


// ------------------------------------------------------------------ main kernel

struct Lane {
    const mbg_desc& D;
    const char* tb;
    u64* lane;
    u32 tid;

    __device__ __forceinline__ u32 xbit(u32 j) const { return (u32)(LANE_X(j >> 6) >> (j & 63)) & 1u; }

    __device__ __forceinline__ u32 key_of(const mbg_nd& n, u32 i) const {
#if MBG_MAX_IN <= 6
        // branch-free: reads past a node's inputs hit the next node's entries (or the
        // table's padding); their bits are masked off once at the end.  Every input id, then
        // every state word, is read before any is used, so the LDS reads overlap
        u32 id[MBG_MAX_IN];
        u64 xw[MBG_MAX_IN];
#pragma unroll
        for (u32 m = 0; m < MBG_MAX_IN; m++) id[m] = TIN(n.in_off + m);
        const u64 own = LANE_X(i >> 6);
#pragma unroll
        for (u32 m = 0; m < MBG_MAX_IN; m++) xw[m] = LANE_X(id[m] >> 6);
        u32 key = (u32)(own >> (i & 63)) & 1u;
#pragma unroll
        for (u32 m = 0; m < MBG_MAX_IN; m++) key |= ((u32)(xw[m] >> (id[m] & 63)) & 1u) << (m + 1);
        key &= (2u << IN_CNT(n)) - 1u;
#else
        u32 key = xbit(i);
        for (u32 m = 0; m < IN_CNT(n); m++) key |= xbit(TIN(n.in_off + m)) << (m + 1);
#endif
        return key;
    }

    // Logic value of fast node n at key (truth table), or of its program.
    __device__ __forceinline__ u32 logic_of(const mbg_nd& n, u32 key) const {
#if MBG_PROG
        if (n.in_cnt == PROG_IN) return eval_prog(D, tb, n.tab_off, [&](u32 w) { return LANE_X(w); });
#endif
        const u32 bp = n.tab_off + key;
        return (TLUT(bp >> 5) >> (bp & 31)) & 1u;
    }

    __device__ __forceinline__ u64 term(const mbg_nd& n, u32 ro, u32 xi, u64 r) const {
        return entropy_term(D, tb, n, ro, xi, r);
    }
};

#if MBG_GENCODE
// Logic value of fast node i evaluated by code generated for the model (a switch on the node
// id, appended to this source), instead of its truth table
__device__ u32 mbg_gen_logic(u32 i, const Lane& L);
#endif

#if MBG_SPARSE
#define MBG_EMPTY_CELL (~0ull)
#define MBG_CELL_PROBES 64
// Adds v to the cell of `key` in the open-addressing table: linear probing from a 64-bit mix
// of the key (the MurmurHash3 finalizer), at most MBG_CELL_PROBES slots.  A slot's key is
// written once (compare-and-swap on empty) and never moves, so no lane waits on another; a
// key that finds no slot is counted in cell_miss and the host repeats the run with a larger
// table (the run is deterministic, so the repeat gives the same cells).
__device__ __forceinline__ void sparse_add(const mbg_desc& D, u64 key, u64 v, u64 q = 0) {
    u64 h = key;
    h ^= h >> 33;
    h *= 0xff51afd7ed558ccdull;
    h ^= h >> 33;
    h *= 0xc4ceb9fe1a85ec53ull;
    h ^= h >> 33;
    for (int probe = 0; probe < MBG_CELL_PROBES; probe++) {
        const u64 slot = (h + (u64)probe) & D.cell_mask;
        u64 cur = __hip_atomic_load(D.cells + slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (cur == MBG_EMPTY_CELL) {
            cur = atomicCAS(D.cells + slot, MBG_EMPTY_CELL, key);
            if (cur == MBG_EMPTY_CELL) cur = key;
        }
        if (cur == key) {
            atomicAdd(D.cell_vals + slot, v);
#if MBG_ERRORS
            if (q) atomicAdd(D.cell_vals2 + slot, q);
#else
            (void)q;
#endif
            return;
        }
    }
    atomicAdd(D.cell_miss, 1ull);
}
#endif

// part: the block's LDS partial sums (MBG_TH_LDS): th[W], then with MBG_CELLS_LDS the window
// table [W][2^K] and the final states [2^K].
__device__ __forceinline__ void flush_cell(const mbg_desc& D, u64* part, int w, sidx s, u64 v) {
#if MBG_SPARSE
    (void)part;
    if (v) sparse_add(D, ((u64)w << D.K) | s, v);
#elif MBG_CELLS_LDS
    if (v) atomicAdd(part + D.W + ((u64)w << D.K) + s, v);
#else
    (void)part;
    if (v) atomicAdd(D.prob + ((u64)w << D.K) + s, v);
#endif
}

// (u64)rint(x) for 0 <= x < 2^52: x + 2^52 lies in [2^52, 2^53), where the spacing of doubles
// is 1, so the addition rounds x to an integer (to nearest, ties to even, as rint does) and
// the low 52 bits of the sum are that integer: one FP64 add and one 32-bit subtract instead of
// rint and a double-to-u64 conversion.  `small` (wave-uniform) says every value of the run
// stays below 2^52; otherwise the general conversion is used.
__device__ __forceinline__ u64 rint_u64(double x, bool small) {
    if (small) return __builtin_bit_cast(u64, x + 4503599627370496.0) - 0x4330000000000000ull;
    return (u64)rint(x);
}

#define DEP_BATCH MBG_DEP_BATCH

// Simulates trajectories of one model (descriptor D) on this wave until its work queue is
// empty and every lane of the wave is idle.
// Window entropy flush: every trajectory adds to th[w] for every window it visits, so the W
// cells are the hottest addresses of the run; with MBG_TH_LDS the block sums them in LDS
// (thl, W entries) and flushes once at its end (exact: integer sums in any order).
__device__ __forceinline__ void flush_th(const mbg_desc& D, u64* thl, int w, u64 v) {
#if MBG_TH_LDS
    (void)D;
    atomicAdd(thl + w, v);
#else
    (void)thl;
    atomicAdd(D.th + w, v);
#endif
}

#if MBG_ERRORS
// Square of a trajectory's window total v in the units of its sum of squares (the oracle's
// sq_units): rint((v * scale)^2 * 2^F).
__device__ __forceinline__ u64 sq_units(u64 v, double scale, int F) {
    const double x = (double)v * scale;
    return (u64)rint(ldexp(x * x, F));
}

// A visit of projected state s lasting v ends inside the current window: its time joins the
// lane's list of the window's states (a trajectory may return to a state within a window, and
// the error columns square the window's total per state).  ne: entries in the list.
__device__ __forceinline__ void err_visit(const mbg_desc& D, u32 t, u32& ne, sidx s, u64 v) {
    if (!v) return;
    for (u32 j = 0; j < ne; j++) {
        const u64 at = (u64)j * D.err_stride + t;
        if (D.err_s[at] == (u64)s) {
            D.err_v[at] += v;
            return;
        }
    }
    if (ne < D.err_cap) {
        const u64 at = (u64)ne * D.err_stride + t;
        D.err_s[at] = (u64)s;
        D.err_v[at] = v;
        ne++;
    } else {
        atomicAdd(D.err_miss, 1ull);  // the host repeats the run with a longer list
    }
}

// The window w ends: every listed state's total and its square go to the tables, and the
// window's TH total th and its square (squares of TH: th2, or after the entropies in LDS).
__device__ __forceinline__ void err_window(const mbg_desc& D, u64* thl, u32 t, u32& ne, int w, u64 th) {
    for (u32 j = 0; j < ne; j++) {
        const u64 at = (u64)j * D.err_stride + t;
        const sidx s = (sidx)D.err_s[at];
        const u64 v = D.err_v[at];
        const u64 q = sq_units(v, D.sq_scale, D.FQ);
#if MBG_SPARSE
        sparse_add(D, ((u64)w << D.K) | s, v, q);
#else
        atomicAdd(D.prob + ((u64)w << D.K) + s, v);
        if (q) atomicAdd(D.prob2 + ((u64)w << D.K) + s, q);
#endif
    }
    ne = 0;
    if (th) {
        flush_th(D, thl, w, th);
        const u64 q = sq_units(th, D.sq_scale_th, D.FQT);
#if MBG_TH_LDS
        if (q) atomicAdd(thl + D.W + w, q);
#else
        if (q) atomicAdd(D.th2 + w, q);
#endif
    }
}
// the current visit (acc_s, acc_p) ends; the current window cur_w ends
#define END_VISIT() err_visit(D, etid, ne, acc_s, acc_p)
#define END_WINDOW()                                   \
    do {                                               \
        err_visit(D, etid, ne, acc_s, acc_p);          \
        err_window(D, thl, etid, ne, cur_w, acc_th);   \
    } while (0)
#else
#define END_VISIT() flush_cell(D, thl, cur_w, acc_s, acc_p)
#define END_WINDOW()                                        \
    do {                                                    \
        flush_cell(D, thl, cur_w, acc_s, acc_p);            \
        if (acc_th) flush_th(D, thl, cur_w, acc_th);        \
    } while (0)
#endif

__device__ __forceinline__ void simulate_variant(const mbg_desc& D, const char* tb, u64* lane, u64* thl) {
    // index of this lane's state: its thread in the block (LDS) or in the grid (HBM)
    const u32 tid = MBG_STATE_HBM == 1 ? blockIdx.x * MBG_BLOCK + threadIdx.x : threadIdx.x;
#if MBG_STATE_HBM >= 2
    const u32 atid = blockIdx.x * MBG_BLOCK + threadIdx.x;  // this lane's column of act in HBM
#endif
    const u32 lid = threadIdx.x & 63u;
#if MBG_ERRORS
    const u32 etid = blockIdx.x * MBG_BLOCK + threadIdx.x;  // this lane's column of the error lists
    u32 ne = 0;
#endif
    Lane L{D, tb, lane, tid};
    const u32 key0 = (u32)D.seed, key1 = (u32)(D.seed >> 32);
    // every slice's fixed-point time and TH x time stays below 2^51 (a slice is at most one
    // window, TH at most log2 63 < 8), so their rounding takes the two-instruction path
#if MBG_DISCRETE
    const bool acc52 = ldexp(8.0, D.FT) < 2251799813685248.0;
#else
    const bool acc52 = ldexp(D.tick, D.FP) < 2251799813685248.0 && ldexp(D.tick * 8.0, D.FT) < 2251799813685248.0;
#endif

    bool have = false, exhausted = false;
    // Step parity is kept equal across the wave: a trajectory starts only on an even loop
    // iteration (a lane refilled on an odd one holds for one iteration), so every active lane
    // has the same k & 1 and the Philox call (one per two steps) runs on every other
    // iteration only, instead of on every iteration for a mix of parities.  (A lane whose
    // dependents take extra iterations steps out of phase: making it wait for its phase
    // measured slower than the extra Philox calls.)
    u32 iter = 0;
    bool hold = false;
    u64 id = 0, total = 0, k = 0, acc_p = 0, acc_th = 0, flips = 0;
    u64 ro_R = 0, ro_S = 0;  // readout rate sum and entropy-term sum (exact)
    u32 pend = 0, pend_off = 0;  // dependents of the last flip not yet re-evaluated
#if MBG_WSUM_REG
    u64 ws[NSUM];
#pragma unroll
    for (int g = 0; g < NSUM; g++) ws[g] = 0;
#endif
    int ro_act = 0;          // readouts with a non-zero rate
    double t = 0.0, th = 0.0;
    sidx s = 0, acc_s = 0;
    int cur_w = -1;
    bool dirty = true;
    u32 d4[4] = {0u, 0u, 0u, 0u};
#if MBG_TRACE
    u64 hash = 0;
#endif

    for (;;) {
        const bool odd_iter = (iter & 1u) != 0u;
        iter++;
        // ---- refill idle lanes from the work queue (one atomic per wave) ----
        const bool need = !have && !exhausted;
        const u64 mask = __ballot(need);
        if (mask) {
            const u32 leader = (u32)__ffsll((long long)mask) - 1u;
            u64 base = 0;
            if (lid == leader) base = atomicAdd(D.work_counter, (u64)__popcll(mask));
            base = __shfl(base, (int)leader);
            if (need) {
                const u64 g = base + (u64)__popcll(mask & ((1ull << lid) - 1ull));
                if (g < D.batch) {
                    have = true;
                    hold = odd_iter;
                    id = D.batch_first + g;
                    // the staged rows are not written while this kernel runs: read through the
                    // constant address space, their loads need not wait for the lane-state
                    // stores in between (which the compiler cannot otherwise tell apart)
                    const CONST_AS u64* st = (const CONST_AS u64*)(D.stage + g * STAGE_ROW);
                    total = 0;
                    // wide states one group at a time, so that the copies of a whole state do
                    // not all sit in registers at once; a group's words are all read before
                    // any is written, so the reads overlap
#if MBG_SUMW > 2
#pragma unroll 1
#else
#pragma unroll
#endif
                    for (int g = 0; g < NSUM; g++) {
                        u64 vx[MBG_SUMW], va[MBG_SUMW], vw[MBG_SUMW];
#pragma unroll
                        for (int q = 0; q < MBG_SUMW; q++) {
                            const int w = g * MBG_SUMW + q;
                            if (MBG_WSUM_REG == 2 && w >= MBG_NW) continue;  // short last group
                            vx[q] = st[w];
                            va[q] = st[MBG_NW + w];
                            vw[q] = st[2 * MBG_NW + w];
                        }
                        u64 gs = 0;
#pragma unroll
                        for (int q = 0; q < MBG_SUMW; q++) {
                            const int w = g * MBG_SUMW + q;
                            if (MBG_WSUM_REG == 2 && w >= MBG_NW) continue;
                            LANE_X(w) = vx[q];
                            LANE_A(w) = va[q];
                            gs += vw[q];
#if MBG_WSUM_REG == 2
                            LANE_W(w) = vw[q];
#endif
                        }
                        LANE_S(g) = gs;
                        total += gs;
                    }
                    s = (sidx)st[3 * MBG_NW + 0];
                    ro_R = st[3 * MBG_NW + 1];
                    ro_S = st[3 * MBG_NW + 2];
                    ro_act = (int)st[3 * MBG_NW + 3];
                    t = 0.0;
                    k = 0;
                    cur_w = -1;
                    acc_p = 0;
                    acc_th = 0;
                    acc_s = s;
                    dirty = true;
                    flips = 0;
#if MBG_ERRORS
                    ne = 0;
#endif
#if MBG_TRACE
                    hash = 0xcbf29ce484222325ull;
#endif
                } else {
                    exhausted = true;
                }
            }
        }
        if (__ballot(have) == 0ull) break;
        if (!have) continue;

        if (hold) {
            hold = false;
            continue;
        }

        // ---- re-evaluate the dependents of the last flip (the flipped node too when it reads
        // itself or is generic), at most MBG_DEP_CAP of them per iteration (0: all): a lane with
        // many dependents then holds the wave for fewer rounds and takes its next step later ----
        if (pend) {
#if MBG_DEP_CAP
            const u32 nq = pend > (u32)MBG_DEP_CAP ? (u32)MBG_DEP_CAP : pend;
#else
            const u32 nq = pend;
#endif
            for (u32 q0 = 0; q0 < nq; q0 += DEP_BATCH) {
                u32 kk[DEP_BATCH], xo[DEP_BATCH], xn[DEP_BATCH];
#if MBG_STATE_HBM >= 2
                u64 aw[DEP_BATCH];
#endif
                u64 rn[DEP_BATCH], rp[DEP_BATCH];
                mbg_nd nk[DEP_BATCH];
#pragma unroll
                for (int u = 0; u < DEP_BATCH; u++) {
                    const bool ok = q0 + u < nq;
                    const u32 dp = TDEP(pend_off + (ok ? q0 + u : 0u));
                    kk[u] = dp & 0xffffffu;
                    const u32 pos = dp >> 24;
#if MBG_STATE_HBM >= 2
                    // the act word is read from HBM before the key gather, whose LDS reads
                    // and logic then cover its latency
                    aw[u] = LANE_A(kk[u] >> 6);
#endif
                    nk[u] = TNODE(kk[u]);
#if MBG_GENCODE
                    const u32 key = L.xbit(kk[u]);  // own state: the generated logic reads the inputs
#else
                    const u32 key = L.key_of(nk[u], kk[u]);
#endif
                    xn[u] = key & 1u;
                    xo[u] = pos == 0 ? xn[u] ^ 1u : xn[u];
#if MBG_GENERIC
                    if (nk[u].flags & F_GEN) {
                        rn[u] = D.rtab[nk[u].tab_off + key];
                        rp[u] = D.rtab[nk[u].tab_off + (key ^ (1u << pos))];
                    } else
#endif
                    {
                        // both table reads unconditional (valid entries), then masked
#if MBG_GENCODE
                        const u32 logic = mbg_gen_logic(kk[u], L);
#else
                        const u32 logic = L.logic_of(nk[u], key);
#endif
#if MBG_STATE_HBM >= 2
                        const u32 a_old = (u32)(aw[u] >> (kk[u] & 63)) & 1u;
#else
                        const u32 a_old = (u32)(LANE_A(kk[u] >> 6) >> (kk[u] & 63)) & 1u;
#endif
                        const u64 r_new = TRATE(2 * kk[u] + xn[u]), r_old = TRATE(2 * kk[u] + xo[u]);
                        rn[u] = xn[u] != logic ? r_new : 0ull;
                        rp[u] = a_old ? r_old : 0ull;
                    }
                    if (!ok) rn[u] = rp[u];
                }
#pragma unroll
                for (int u = 0; u < DEP_BATCH; u++) {
                    if (rn[u] != rp[u]) {
                        const u32 wk = kk[u] >> 6;
                        const u64 bit = 1ull << (kk[u] & 63);
                        WS_ADD(wk, rn[u] - rp[u]);
                        total += rn[u] - rp[u];
#if MBG_STATE_HBM >= 2
                        // act words of two dependents in one batch may coincide: the batch's
                        // earlier writes are in memory, so read the word again only then
                        u64 a = aw[u];
                        _Pragma("unroll") for (int v = 0; v < u; v++) if ((kk[v] >> 6) == wk) a = LANE_A(wk);
                        const u64 an = rn[u] ? (a | bit) : (a & ~bit);
                        if (an != a) LANE_A(wk) = an;
#else
                        const u64 a = LANE_A(wk);
                        LANE_A(wk) = rn[u] ? (a | bit) : (a & ~bit);
#endif
                        if (nk[u].flags & F_RO) {
                            const u32 ro = (nk[u].flags & F_RO) - 1u;
                            ro_R += rn[u] - rp[u];
                            ro_S += L.term(nk[u], ro, xn[u], rn[u]) - L.term(nk[u], ro, xo[u], rp[u]);
                            ro_act += (int)(rn[u] != 0) - (int)(rp[u] != 0);
                            dirty = true;
                        }
                    }
                }
            }
            pend -= nq;
            pend_off += nq;
            if (pend) continue;
        }

        // ---- one Gillespie step (src/jit_kernels/simulation.cu:106-148) ----
        double t1, th_use = 0.0;
        u32 dr2 = 0;
        const bool fixed_pt = total == 0;
        if (fixed_pt) {
            t1 = D.maxt;
        } else {
            if ((k & 1ull) == 0ull) {
                // the key schedule is recomputed per call on the scalar unit (the seed is
                // wave-uniform): hoisted out of the loop, its 20 round keys would not fit in
                // SGPRs and would come back through VALU lane reads
                u32 k0 = __builtin_amdgcn_readfirstlane(key0), k1 = __builtin_amdgcn_readfirstlane(key1);
                MBG_OPAQUE_S(k0);
                MBG_OPAQUE_S(k1);
                mbg_philox((u32)(k >> 1), 1u, (u32)id, D.variant_id, k0, k1, d4);
            }
            const u32 dr1 = (k & 1ull) ? d4[2] : d4[0];
            dr2 = (k & 1ull) ? d4[3] : d4[1];
#if MBG_DISCRETE
            (void)dr1;
            t1 = (k + 1 >= D.nsteps) ? D.maxt : (double)(k + 1) * D.tick;
#else
            t1 = t + mbg_neglog_u32(dr1) / ((double)total * D.rscale);
#endif
            if (t1 > D.maxt) t1 = D.maxt;
            if (dirty) {
                // Shannon entropy (base 2) of the readout rates (src/generator.cpp:146-175):
                // TH = log2 R - S 2^-sigma / R, 0 with fewer than two non-zero readout rates
                th = 0.0;
                if (ro_act >= 2) {
                    const double Rd = (double)ro_R;
                    th = mbg_log2(Rd) - ldexp((double)ro_S, -D.sigma) / Rd;
                    if (th < 0.0) th = 0.0;
                }
                dirty = false;
            }
            th_use = th;
        }

        // ---- statistics of the sojourn [t, t1) in projected state s ----
#if MBG_DISCRETE
        {
            int w = (int)floor(t * D.inv_tick + 0.5);
            if (w > D.W - 1) w = D.W - 1;
            if (w != cur_w) {
                if (cur_w >= 0) END_WINDOW();
                cur_w = w;
                acc_p = 0;
                acc_th = 0;
                acc_s = s;
            } else if (s != acc_s) {
                END_VISIT();
                acc_p = 0;
                acc_s = s;
            }
            acc_p += 1;
            acc_th += rint_u64(ldexp(th_use, D.FT), acc52);
        }
#else
        {
            int w = (int)floor(t * D.inv_tick);
            double b = t;
            while (t1 > b) {
                const double we = (double)(w + 1) * D.tick;
                const double e = t1 < we ? t1 : we;
                const double sl = e - b;
                const int ww = w > D.W - 1 ? D.W - 1 : w;
                if (ww != cur_w) {
                    if (cur_w >= 0) END_WINDOW();
                    cur_w = ww;
                    acc_p = 0;
                    acc_th = 0;
                    acc_s = s;
                } else if (s != acc_s) {
                    END_VISIT();
                    acc_p = 0;
                    acc_s = s;
                }
                acc_p += rint_u64(ldexp(sl, D.FP), acc52);
                acc_th += rint_u64(ldexp(th_use * sl, D.FT), acc52);
                w++;
                b = e;
            }
        }
#endif
        t = t1;

        if (t >= D.maxt) {
            // ---- trajectory end: flush, final state, fixed point, summary ----
            if (cur_w >= 0) END_WINDOW();
#if MBG_SPARSE
            sparse_add(D, ((u64)D.W << D.K) | s, 1ull);
#elif MBG_CELLS_LDS
            atomicAdd(thl + D.W + ((u64)D.W << D.K) + s, 1ull);
#else
            atomicAdd(D.fin + s, 1ull);
#endif
            atomicAdd(D.flip_total, flips);
            if (fixed_pt) {
                const u64 slot = atomicAdd(D.fp_count, 1ull);
                if (slot < D.fp_cap) {
#pragma unroll 4
                    for (int w = 0; w < MBG_NW; w++) D.fp_states[slot * MBG_NW + w] = LANE_X(w);
                }
            }
#if MBG_TRACE
            const u64 loc = id - D.first_id;
            if (loc < D.summ_cap) {
                D.summ[4 * loc + 0] = flips;
                D.summ[4 * loc + 1] = hash;
                D.summ[4 * loc + 2] = __builtin_bit_cast(u64, t);
                D.summ[4 * loc + 3] = fixed_pt ? 2ull : 1ull;
            }
#endif
            have = false;
            continue;
        }
        // compatibility with the reference's 100-record launches: the 99th step of a launch is
        // recorded but not followed by a flip, and the next launch draws a fresh waiting time for
        // the same state (src/jit_kernels/simulation.cu:104-147)
        if (D.compat && (k + 1ull) % 99ull == 0ull) {
            k++;
            continue;
        }

        // ---- flip selection: group sums, then the active bits of the group's words ----
        const u64 target = mbg_flip_target(dr2, total);
        u64 acc = 0, sgs = 1;
        int gsel = NSUM - 1;
        bool found = false;
        // branch-free: selects on a lane mask instead of a branch per group
#pragma unroll
        for (int g = 0; g < NSUM; g++) {
            const u64 sg = LANE_S(g);
            const u64 na = acc + sg;
            const bool hit = !found && na > target;
            gsel = hit ? g : gsel;
            sgs = hit ? sg : sgs;
            acc = (found || hit) ? acc : na;
            found = found || hit;
        }
        // In the group, the selected node is the first (in node order) whose inclusive prefix
        // of rates exceeds rem = target - acc.  Equivalently it is the first from the top whose
        // inclusive suffix reaches sgs - rem (the group sum minus rem): each lane searches from
        // the nearer end, which halves the rounds the slowest lane of the wave needs.  Both
        // directions share one loop: `lim` is rem from the bottom (hit when rate > lim) and
        // sgs - rem - 1 from the top (hit when rate >= sgs - rem), and from the top the active
        // bits are read bit-reversed so that the lowest set bit is always the next candidate.
#if MBG_WSUM_REG == 2
        // second level: the word sums of the selected group (in LDS, or in HBM), all read
        // before the first is used so that the reads overlap, then selected branch-free
        u32 w0 = (u32)(gsel * MBG_SUMW);
        {
            bool f2 = false;
            u64 sw_sel = 1;
            u32 wsel = w0;
            u64 swq[MBG_SUMW];
#pragma unroll
            for (int q = 0; q < MBG_SUMW; q++) {
                const u32 w = (u32)(gsel * MBG_SUMW + q);
                swq[q] = w < (u32)MBG_NW ? LANE_W(w) : 0ull;
            }
#pragma unroll
            for (int q = 0; q < MBG_SUMW; q++) {
                const u32 w = (u32)(gsel * MBG_SUMW + q);
                const bool h2 = !f2 && w < (u32)MBG_NW && acc + swq[q] > target;
                wsel = h2 ? w : wsel;
                sw_sel = h2 ? swq[q] : sw_sel;
                acc = (f2 || h2 || w >= (u32)MBG_NW) ? acc : acc + swq[q];
                f2 = f2 || h2;
            }
            found = found && f2;
            sgs = sw_sel;
            w0 = wsel;
        }
#define SEARCH_W 1
#else
        const u32 w0 = (u32)(gsel * MBG_SUMW);
#define SEARCH_W MBG_SUMW
#endif
        const u64 rem0 = target - acc;
        const bool down = found && rem0 >= (sgs >> 1);
        u64 lim = down ? sgs - rem0 - 1ull : rem0;
        // guard for an inconsistent group (never taken when the sums are exact)
        u32 j = w0 * 64u;
        u64 r_sel = 0, xw = LANE_X(j >> 6), a_hit = 0;
        bool hit = false;
        for (int q0 = 0; q0 < SEARCH_W && !hit; q0++) {
            const int q = down ? SEARCH_W - 1 - q0 : q0;
            const u32 wi = w0 + (u32)q;
            const u64 am = LANE_A(wi);
            u64 m = down ? (((u64)__builtin_bitreverse32((u32)am) << 32) | (u64)__builtin_bitreverse32((u32)(am >> 32)))
                         : am;
            const u64 xq = LANE_X(wi);
#if MBG_GENERIC
            const u64 gm = TGEN(wi);
#endif
            const u32 wbase = wi * 64u;
            const u32 flip = down ? 63u : 0u;  // bit-reversed order: candidate t is bit 63 - t
            while (m) {
                // MBG_SCAN candidates per round trip to the rate table; past the last active
                // bit (t = 64) a candidate still reads a valid entry (bit 0 or 63 of the word)
                // and its rate is masked to 0
                u32 bb[MBG_SCAN];
                u64 rr[MBG_SCAN], mm = m;
#pragma unroll
                for (int c = 0; c < MBG_SCAN; c++) {
                    const u32 t = mm ? (u32)__builtin_ctzll(mm) : 64u;
                    const u32 b = (t ^ flip) & 63u;
                    bb[c] = b;
                    mm &= mm - 1ull;
                    const u64 r_ = TRATE(2 * (wbase + b) + ((xq >> b) & 1ull));
                    rr[c] = t < 64u ? r_ : 0ull;
#if MBG_GENERIC
                    if (t < 64u && ((gm >> b) & 1ull)) {
                        const mbg_nd nd = TNODE(wbase + b);
                        rr[c] = rate_from_key(D, tb, nd, wbase + b, L.key_of(nd, wbase + b));
                    }
#endif
                }
#pragma unroll
                for (int c = 0; c < MBG_SCAN; c++) {
                    // branch-free: a candidate past the word's last active bit has rate 0 and
                    // is never a hit (lim < 0 is false), so a hit's bb[c] is a real candidate
                    const bool h = !hit && lim < rr[c];
                    j = h ? wbase + bb[c] : j;
                    r_sel = h ? rr[c] : r_sel;
                    lim = hit || h ? lim : lim - rr[c];
                    hit = hit || h;
                }
                if (hit) {
                    xw = xq;
                    a_hit = am;
                    break;
                }
                m = mm;
            }
        }

        // ---- flip j and update the exact sums ----
        const u32 xj_old = (u32)(xw >> (j & 63)) & 1u;
        LANE_X(j >> 6) = xw ^ (1ull << (j & 63));
        const mbg_nd nj = TNODE(j);
        if (!(nj.flags & F_SELF)) {
            // fast node that does not read itself: it was active, it now agrees with its
            // logic, so its rate drops from r_sel to 0
            const u32 wj = j >> 6;
            WS_ADD(wj, 0ull - r_sel);
            total -= r_sel;
            // the hit's act word was read by the search (HBM act: no second read)
            LANE_A(wj) = (MBG_STATE_HBM >= 2 && hit ? a_hit : LANE_A(wj)) & ~(1ull << (j & 63));
            if (nj.flags & F_RO) {
                ro_R -= r_sel;
                ro_S -= L.term(nj, (nj.flags & F_RO) - 1u, xj_old, r_sel);
                ro_act -= 1;
                dirty = true;
            }
        }
        // its dependents are re-evaluated from the next iteration on
        pend = nj.dep_cnt;
        pend_off = nj.dep_off;
        if (nj.flags & F_RO) s ^= (sidx)1 << ((nj.flags & F_RO) - 1u);
        flips++;
#if MBG_TRACE
        hash = (hash ^ (u64)j) * 0x100000001b3ull;
#endif
        k++;
    }
}

#if !MBG_MULTI
extern "C" __global__ __launch_bounds__(MBG_BLOCK) void mbg_simulate(mbg_desc D) {
    MBG_DYN_LDS(smem);
#if MBG_TABLES_LDS
    // stage the model tables in LDS once per (persistent) block
    {
        const u32 words = D.lay.bytes / 8;
        const u64* src = (const u64*)D.blob;
        for (u32 i = threadIdx.x; i < words; i += MBG_BLOCK) smem[i] = src[i];
        __syncthreads();
    }
    const char* tb = (const char*)smem;
    u64* after = smem + D.lay.bytes / 8;
#else
    const char* tb = D.blob;
    u64* after = smem;
#endif
#if MBG_STATE_HBM == 1
    u64* lane = D.lane_state;
    u64* part = after;
#else
    u64* lane = after;
    u64* part = lane + (MBG_STATE_HBM >= 2 ? 1 : MBG_WSUM_REG == 1 ? 2 : 3) * MBG_NW * MBG_BLOCK;
#endif
#if MBG_TH_LDS
    // the block's partial sums follow the lanes' state (or the tables): th[W], then
    // (MBG_CELLS_LDS) the window table [W][2^K] and the final states [2^K]
    u64* thl = part;
#if MBG_CELLS_LDS
    const u32 S = 1u << D.K;
    const u32 nparts = (u32)D.W + ((u32)D.W + 1u) * S;
#elif MBG_ERRORS
    const u32 nparts = 2u * (u32)D.W;  // th[W], then the squares th2[W]
#else
    const u32 nparts = (u32)D.W;
#endif
    for (u32 i = threadIdx.x; i < nparts; i += MBG_BLOCK) thl[i] = 0;
    __syncthreads();
    simulate_variant(D, tb, lane, thl);
    __syncthreads();  // every wave of the block has left its loop
    for (u32 i = threadIdx.x; i < nparts; i += MBG_BLOCK) {
        const u64 v = thl[i];
        if (!v) continue;
#if MBG_CELLS_LDS
        const u32 c = i - (u32)D.W;  // cell index past the entropies
        if (i < (u32)D.W) atomicAdd(D.th + i, v);
        else if (c < (u32)D.W * S) atomicAdd(D.prob + c, v);
        else atomicAdd(D.fin + (c - (u32)D.W * S), v);
#elif MBG_ERRORS
        if (i < (u32)D.W) atomicAdd(D.th + i, v);
        else atomicAdd(D.th2 + (i - (u32)D.W), v);
#else
        atomicAdd(D.th + i, v);
#endif
    }
#else
    (void)part;
    simulate_variant(D, tb, lane, nullptr);
#endif
}
#else
// Variant sweep: nv models of one state-word class in one persistent launch.  Each wave
// drains one variant at a time (its descriptor is wave-uniform, read with scalar loads) and
// moves on when that variant's queue is empty.
// Tables in HBM (MBG_TABLES_LDS 0): the waves start at a variant chosen by MBG_SWEEP_XCD.
// Tables in LDS (MBG_TABLES_LDS 1): every wave of a block works on the block's current variant,
// whose tables the block stages in the first tab_words words of its LDS (the class's largest
// table blob); a block starts at variant block_start[block] (a share of the class's work
// proportional to the variants' work, or variant 0 when block_start is null) and moves on, all
// its waves together, when that variant's queue is empty.
// With MBG_TH_LDS each wave sums its current variant's window entropies in its own LDS slice
// of wmax cells (after the lanes' state) and flushes them when it moves to the next variant.
// clk (may be null): the class's first block start and last block end on the device's
// constant-rate clock, by atomic min / max (the class's kernel time inside a graph).
extern "C" __global__ __launch_bounds__(MBG_BLOCK) void mbg_simulate_multi(const CONST_AS mbg_desc* descs, int nv, int wmax,
                                                                           u64* clk, int tab_words,
                                                                           const int* block_start) {
    MBG_DYN_LDS(smem);
    if (clk && threadIdx.x == 0) atomicMin(clk, (u64)wall_clock64());
#if MBG_TABLES_LDS
    u64* ctl = smem + tab_words;  // block-uniform flag: the current variant still has work
    u64* base = ctl + 2;          // the lanes' state (16-byte aligned after the tables)
    const int start = block_start ? block_start[blockIdx.x] : 0;
#else
    (void)tab_words;
    (void)block_start;
    u64* base = smem;
#if MBG_SWEEP_XCD == 2
    // every wave starts at the class's first variant: the host orders them by decreasing work
    // per trajectory, so the costliest trajectories start first and the run ends on the cheapest
    const int start = 0;
#elif MBG_SWEEP_XCD
    // the waves of a block start at one variant, and the blocks of one XCD (blocks are dispatched
    // to the 8 XCDs in turn) at consecutive variants of one eighth of the class: an XCD's L2 then
    // holds the tables of the variants its CUs work on, and a CU's waves share their L1 lines
    const u32 per_xcd = (gridDim.x + 7u) / 8u;
    const u32 order = (blockIdx.x % 8u) * per_xcd + blockIdx.x / 8u;
    const int start = (int)(((u64)order * (u64)nv) / (8ull * per_xcd));
#else
    const u32 waves_total = gridDim.x * (MBG_BLOCK / 64);
    const u32 wave = blockIdx.x * (MBG_BLOCK / 64) + (threadIdx.x >> 6);
    const int start = (int)(((u64)wave * (u64)nv) / waves_total);
#endif
#endif
#if MBG_TH_LDS
    // the slice holds the variant's W entropies, then (MBG_ERRORS) their W squares
    const u32 lid = threadIdx.x & 63u;
    const u32 th_cells = (MBG_ERRORS ? 2u : 1u);
    u64* thw = base + (MBG_STATE_HBM == 1 ? 0 : MBG_STATE_HBM == 2 ? 1 : MBG_WSUM_REG == 1 ? 2 : 3) * MBG_NW * MBG_BLOCK +
               (threadIdx.x >> 6) * th_cells * (u32)wmax;
#else
    (void)wmax;
    u64* thw = nullptr;
#endif
    for (int i = 0; i < nv; i++) {
        int v = start + i;
        if (v >= nv) v -= nv;
        v = __builtin_amdgcn_readfirstlane(v);
        // the descriptors are constant during the launch: read through the constant address
        // space, their fields are scalar loads the compiler may repeat where it needs them,
        // instead of vector loads kept in VGPRs or reloaded and waited for inside the step loop
        const mbg_desc& D = *(const mbg_desc*)(descs + v);
#if MBG_TABLES_LDS
        __syncthreads();  // every wave has left the previous variant: the tables and the flag are free
        if (threadIdx.x == 0)
            ctl[0] = __hip_atomic_load(D.work_counter, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < D.batch ? 1ull : 0ull;
        __syncthreads();
        if (ctl[0] == 0ull) continue;  // block-uniform: the variant's queue is already empty
        {
            const u32 words = D.lay.bytes / 8;
            const u64* src = (const u64*)D.blob;
            for (u32 k = threadIdx.x; k < words; k += MBG_BLOCK) smem[k] = src[k];
        }
        __syncthreads();
        const char* tb = (const char*)smem;
#else
        const char* tb = D.blob;
#endif
#if MBG_TH_LDS
        for (u32 c = lid; c < th_cells * (u32)D.W; c += 64u) thw[c] = 0;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
#endif
        // the lanes' state in LDS, or with MBG_STATE_HBM 1 in the class's HBM buffer (every
        // descriptor of the class names the same one)
        simulate_variant(D, tb, MBG_STATE_HBM == 1 ? D.lane_state : base, thw);
#if MBG_TH_LDS
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        __builtin_amdgcn_wave_barrier();  // every lane of the wave has left the variant's loop
        for (u32 c = lid; c < th_cells * (u32)D.W; c += 64u) {
            const u64 x = thw[c];
            if (!x) continue;
            if (c < (u32)D.W) atomicAdd(D.th + c, x);
            else atomicAdd(D.th2 + (c - (u32)D.W), x);
        }
#endif
    }
    if (clk) {
        __syncthreads();
        if (threadIdx.x == 0) atomicMax(clk + 1, (u64)wall_clock64());
    }
}
#endif

// ------------------------------------------------------------------ readout marginals

// Marginal time each readout is ON per window, from a window table: out[w][j] = sum over the
// projected states s with bit j set of prob[w][s] (exact 64-bit sums).  One block per window.
__device__ __forceinline__ void marginals_block(const u64* prob, int K, u32 w, u64* out) {
    __shared__ u64 part[256 * 20];
    const u32 tid = threadIdx.x;
    const u64 S = 1ull << K;
    const u64* row = prob + ((u64)w << K);
    u64 acc[20];
#pragma unroll
    for (int j = 0; j < 20; j++) acc[j] = 0;
    for (u64 s = tid; s < S; s += 256) {
        const u64 v = row[s];
#pragma unroll
        for (int j = 0; j < 20; j++) acc[j] += ((s >> j) & 1ull) ? v : 0ull;
    }
#pragma unroll
    for (int j = 0; j < 20; j++) part[j * 256 + tid] = acc[j];
    __syncthreads();
    if (tid < (u32)K) {
        u64 sum = 0;
        for (int i = 0; i < 256; i++) sum += part[tid * 256 + i];
        out[(u64)w * K + tid] = sum;
    }
}

extern "C" __global__ __launch_bounds__(256) void mbg_marginals(const u64* prob, int K, u64* out) {
    marginals_block(prob, K, blockIdx.x, out);
}

// The marginals of many window tables (the variants of a sweep) in one launch: block (w, v)
// reduces window w of job v into out + out_off.
struct mbg_marg_job {
    const u64* prob;
    u64 out_off;
    int W, K;
};

extern "C" __global__ __launch_bounds__(256) void mbg_marginals_batch(const mbg_marg_job* jobs, u64* out) {
    const mbg_marg_job J = jobs[blockIdx.y];
    if ((int)blockIdx.x >= J.W) return;
    marginals_block(J.prob, J.K, blockIdx.x, out + J.out_off);
}

// ------------------------------------------------------------------ fixed-point deduplication

// After a run: count the distinct states among the n appended fixed-point records (rows of
// `stride` words, the first `words` significant) in an open-addressing table of cap_mask + 1
// slots (at least 2n).  A slot holds the index of the first record of its state; records are
// immutable here, so a thread that finds a slot taken compares the two records directly and
// never waits for another thread.  Every record ends in a slot of its own state: an empty slot
// exists since the table is at most half full.
extern "C" __global__ __launch_bounds__(256) void mbg_fp_dedup(const u64* states, u64 n, int stride, int words,
                                                              u32* keys, u64* counts, u64 cap_mask) {
    const u64 i = (u64)blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    const u64* s = states + i * (u64)stride;
    u64 h = 0x9E3779B97F4A7C15ull;
    for (int w = 0; w < words; w++) {
        h ^= s[w];
        h *= 0xBF58476D1CE4E5B9ull;
        h ^= h >> 31;
    }
    u64 slot = h & cap_mask;
    for (;;) {
        const u32 prev = atomicCAS(keys + slot, 0xFFFFFFFFu, (u32)i);
        if (prev == 0xFFFFFFFFu) break;
        const u64* t = states + (u64)prev * (u64)stride;
        bool eq = true;
        for (int w = 0; w < words; w++)
            if (t[w] != s[w]) {
                eq = false;
                break;
            }
        if (eq) break;
        slot = (slot + 1) & cap_mask;
    }
    atomicAdd(counts + slot, 1ull);
}

// Occupied slots to (first record index, count) pairs.
extern "C" __global__ __launch_bounds__(256) void mbg_fp_compact(const u32* keys, const u64* counts, u64 cap, u32* out_idx,
                                                                u64* out_cnt, u64* out_n) {
    const u64 slot = (u64)blockIdx.x * 256u + threadIdx.x;
    if (slot >= cap) return;
    const u32 k = keys[slot];
    if (k == 0xFFFFFFFFu) return;
    const u64 pos = atomicAdd(out_n, 1ull);
    out_idx[pos] = k;
    out_cnt[pos] = counts[slot];
}

// The distinct states: out[j][w] = record idx[j], word w.
extern "C" __global__ __launch_bounds__(256) void mbg_fp_gather(const u64* states, int stride, int words, const u32* idx,
                                                               u64 m, u64* out) {
    const u64 t = (u64)blockIdx.x * 256u + threadIdx.x;
    if (t >= m * (u64)words) return;
    const u64 j = t / (u64)words;
    const int w = (int)(t - j * (u64)words);
    out[t] = states[(u64)idx[j] * (u64)stride + w];
}

// ------------------------------------------------------------------ sparse table read-back
// Taken slots (key != ~0) of the sparse cell table: mbg_cells_count counts them, then
// mbg_cells_compact copies them to dense arrays, in no particular order (the host sorts).
// Both reduce per block (wave ballots, then LDS) and issue one atomic per block.
__device__ __forceinline__ u32 block_exclusive_count(bool has, u32* lds, u32& total) {
    const u32 tid = threadIdx.x, lid = tid & 63u, wid = tid >> 6;
    const u64 mask = __ballot(has);
    if (lid == 0) lds[wid] = (u32)__popcll(mask);
    __syncthreads();
    u32 before = 0;
    total = 0;
    for (u32 k = 0; k < 4u; k++) {
        before += k < wid ? lds[k] : 0u;
        total += lds[k];
    }
    return before + (u32)__popcll(mask & ((1ull << lid) - 1ull));
}

extern "C" __global__ void __launch_bounds__(256) mbg_cells_count(const u64* keys, u64 cap, u64* out_n) {
    __shared__ u32 lds[4];
    const u64 i = (u64)blockIdx.x * 256u + threadIdx.x;
    u32 total;
    (void)block_exclusive_count(i < cap && keys[i] != ~0ull, lds, total);
    if (threadIdx.x == 0 && total) atomicAdd(out_n, (u64)total);
}

// vals2 / out_vals2 (may be null): the sums of squares beside the values (error columns).
extern "C" __global__ void __launch_bounds__(256) mbg_cells_compact(const u64* keys, const u64* vals,
                                                                    const u64* vals2, u64 cap, u64* out_keys,
                                                                    u64* out_vals, u64* out_vals2, u64* out_n,
                                                                    u64 out_cap) {
    __shared__ u32 lds[4];
    __shared__ u64 base;
    const u64 i = (u64)blockIdx.x * 256u + threadIdx.x;
    const bool has = i < cap && keys[i] != ~0ull;
    u32 total;
    const u32 rank = block_exclusive_count(has, lds, total);
    if (threadIdx.x == 0) base = total ? atomicAdd(out_n, (u64)total) : 0ull;
    __syncthreads();
    const u64 pos = base + rank;
    if (has && pos < out_cap) {  // out_n above out_cap is reported to the host as an error
        out_keys[pos] = keys[i];
        out_vals[pos] = vals[i];
        if (vals2) out_vals2[pos] = vals2[i];
    }
}
