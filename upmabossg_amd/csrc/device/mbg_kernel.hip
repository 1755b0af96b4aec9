// Persistent Gillespie trajectory kernel + statistics stage for gfx950.
//
// Compiled at run time by hipRTC.  The code generator prepends, per model:
//   MBG_NW       64-bit state words per trajectory (ceil(N / 64))
//   MBG_BLOCK    threads per block (a multiple of 64)
//   MBG_DISCRETE 1 for discrete_time models
//   MBG_TRACE    1 to record per-trajectory summaries (flip count, flip hash)
// and mbg_math.h.
//
// One trajectory per lane (reference: one CUDA thread per trajectory,
// src/jit_kernels/simulation.cu:77-164).  Unlike the reference's O(N) rate
// recompute and linear flip scan per step (src/generator.cpp:124-144,
// simulation.cu:9-21), each lane keeps
//   x[w]    the Boolean state,
//   act[w]  the active set (nodes with a non-zero rate),
//   wsum[w] the exact fixed-point rate sum of each 64-node word,
// in LDS, lane-interleaved (element w of lane l at w*BLOCK + l, so lanes never
// share a bank whatever word each one touches).  A transition re-evaluates
// only the flipped node and its dependents from per-node truth tables and
// updates the sums with exact integer deltas, so the two-level flip search
// (word sums, then the active bits of one word) selects exactly the node the
// reference's linear scan of prefix sums would.
//
// Statistics (src/jit_kernels/window_average_small.cu:50-136,
// final_states.cu:9-21, src/statistics/fixed_states.cu:84-125) are
// accumulated per lane in registers for the current (window, projected state)
// cell and flushed with 64-bit integer atomics when the cell changes, so the
// tables are exact sums independent of scheduling.

typedef unsigned int u32;
typedef unsigned long long u64;

struct mbg_node {
    u32 in_off;               // first input in inputs[]
    u32 dep_off;              // first dependent in deps[]
    u32 tab_off;              // fast: bit offset into lut[]; generic: index into rtab[]
    unsigned short dep_cnt;   // dependents (nodes whose rate reads this node, itself included)
    unsigned char in_cnt;     // inputs other than the node itself
    unsigned char flags;      // bit 7: generic rate table; bits 0-6: readout index + 1 (0: internal)
};

struct mbg_desc {
    const mbg_node* nodes;
    const u32* inputs;        // node ids
    const u32* deps;          // node id | (key bit position << 24)
    const u32* lut;           // logic truth tables, key = x_i | inputs << 1
    const u64* rates;         // fixed-point [rate_up, rate_down] per node (fast nodes)
    const u64* rtab;          // fixed-point rate per key (generic nodes)
    const u64* gen_mask;      // per word: nodes with a generic rate table
    const float* istate;      // initial ON probability per node
    const u32* readouts;      // non-internal node ids in node order
    const double* ro_lg;      // per readout: U*log2(U), D*log2(D) of the fixed-point rates
    u64* prob;                // [W][2^K] fixed-point sojourn time (counts when discrete)
    u64* th;                  // [W] fixed-point transition entropy * time
    u64* fin;                 // [2^K] final projected states
    u64* fp_states;           // [fp_cap][MBG_NW] fixed-point states, append order
    u64* fp_count;
    u64* summ;                // [summ_cap][4] flips, flip hash, final time bits, status
    u64* work_counter;        // next trajectory to start
    u64* flip_total;          // transitions taken, all trajectories
    u64 fp_cap;
    u64 summ_cap;
    u64 nsteps;               // discrete: steps per trajectory
    u64 n_traj;
    u64 first_id;             // trajectory ids are first_id + work index
    u64 seed;
    double maxt, tick, inv_tick, rscale;
    u32 variant_id;
    int N, K, W, FP, FT;
};

#define LANE_X(w) lane_mem[(0 * MBG_NW + (w)) * MBG_BLOCK + tid]
#define LANE_A(w) lane_mem[(1 * MBG_NW + (w)) * MBG_BLOCK + tid]
#define LANE_S(w) lane_mem[(2 * MBG_NW + (w)) * MBG_BLOCK + tid]

struct Lane {
    const mbg_desc& D;
    u64* lane_mem;
    u32 tid;

    __device__ __forceinline__ u32 xbit(u32 j) const { return (u32)(LANE_X(j >> 6) >> (j & 63)) & 1u; }

    __device__ __forceinline__ u32 key_of(const mbg_node& n, u32 i) const {
        u32 key = xbit(i);
        for (u32 m = 0; m < n.in_cnt; m++) key |= xbit(D.inputs[n.in_off + m]) << (m + 1);
        return key;
    }

    __device__ __forceinline__ u64 rate_of(const mbg_node& n, u32 i, u32 key) const {
        if (n.flags & 0x80) return D.rtab[n.tab_off + key];
        const u32 bp = n.tab_off + key;
        const u32 logic = (D.lut[bp >> 5] >> (bp & 31)) & 1u;
        const u32 xi = key & 1u;
        return xi == logic ? 0ull : D.rates[2 * i + xi];
    }

    // Shannon entropy (base 2) of the readout rates (src/generator.cpp:146-175), FP64.
    __device__ double entropy() const {
        u64 R = 0;
        int act = 0;
        double S = 0.0;
        for (int j = 0; j < D.K; j++) {
            const u32 n = D.readouts[j];
            const mbg_node nd = D.nodes[n];
            u64 r;
            double term;
            if (nd.flags & 0x80) {
                r = rate_of(nd, n, key_of(nd, n));
                const double rd = (double)r;
                term = rd * mbg_log2(rd);
            } else {
                const u32 a = (u32)(LANE_A(n >> 6) >> (n & 63)) & 1u;
                const u32 xi = xbit(n);
                r = a ? D.rates[2 * n + xi] : 0ull;
                term = D.ro_lg[2 * j + xi];
            }
            if (r) {
                R += r;
                act++;
                S = S + term;
            }
        }
        if (act < 2) return 0.0;
        const double Rd = (double)R;
        const double v = mbg_log2(Rd) - S / Rd;
        return v < 0.0 ? 0.0 : v;
    }
};

__device__ __forceinline__ void flush_cell(const mbg_desc& D, int w, u32 s, u64 v) {
    if (v) atomicAdd(D.prob + ((u64)w << D.K) + s, v);
}

extern "C" __global__ __launch_bounds__(MBG_BLOCK) void mbg_simulate(mbg_desc D) {
    __shared__ u64 lane_mem[3 * MBG_NW * MBG_BLOCK];
    const u32 tid = threadIdx.x;
    const u32 lane = tid & 63u;
    Lane L{D, lane_mem, tid};
    const u32 key0 = (u32)D.seed, key1 = (u32)(D.seed >> 32);

    bool have = false, exhausted = false;
    u64 id = 0, total = 0, k = 0, acc_p = 0, acc_th = 0;
    double t = 0.0, th = 0.0;
    u32 s = 0, acc_s = 0;
    int cur_w = -1;
    bool dirty = true;
    u32 d4[4] = {0u, 0u, 0u, 0u};
    u64 flips = 0;
#if MBG_TRACE
    u64 hash = 0;
#endif

    for (;;) {
        // ---- refill idle lanes from the work queue (one atomic per wave) ----
        const bool need = !have && !exhausted;
        const u64 mask = __ballot(need);
        if (mask) {
            const u32 leader = (u32)__ffsll((long long)mask) - 1u;
            u64 base = 0;
            if (lane == leader) base = atomicAdd(D.work_counter, (u64)__popcll(mask));
            base = __shfl(base, (int)leader);
            if (need) {
                const u64 g = base + (u64)__popcll(mask & ((1ull << lane) - 1ull));
                if (g < D.n_traj) {
                    have = true;
                    id = D.first_id + g;
                    // initial state: draw i of the trajectory's stream 0
                    // (src/jit_kernels/simulation.cu:34-71)
                    for (int w = 0; w < MBG_NW; w++) {
                        u64 word = 0;
                        for (int q = 0; q < 16; q++) {
                            const int i0 = w * 64 + q * 4;
                            if (i0 >= D.N) break;
                            u32 o[4];
                            mbg_philox((u32)(i0 >> 2), 0u, (u32)id, D.variant_id, key0, key1, o);
                            for (int r = 0; r < 4; r++) {
                                const int i = i0 + r;
                                if (i < D.N && mbg_unit_f32(o[r]) <= D.istate[i]) word |= 1ull << (i & 63);
                            }
                        }
                        LANE_X(w) = word;
                    }
                    total = 0;
                    for (int w = 0; w < MBG_NW; w++) {
                        u64 a = 0, sum = 0;
                        for (int b = 0; b < 64; b++) {
                            const u32 i = (u32)(w * 64 + b);
                            if ((int)i >= D.N) break;
                            const mbg_node nd = D.nodes[i];
                            const u64 r = L.rate_of(nd, i, L.key_of(nd, i));
                            if (r) {
                                a |= 1ull << b;
                                sum += r;
                            }
                        }
                        LANE_A(w) = a;
                        LANE_S(w) = sum;
                        total += sum;
                    }
                    s = 0;
                    for (int j = 0; j < D.K; j++) s |= L.xbit(D.readouts[j]) << j;
                    t = 0.0;
                    k = 0;
                    cur_w = -1;
                    acc_p = 0;
                    acc_th = 0;
                    acc_s = s;
                    dirty = true;
                    flips = 0;
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

        // ---- one Gillespie step (src/jit_kernels/simulation.cu:106-148) ----
        double t1, th_use = 0.0;
        u32 dr2 = 0;
        const bool fixed_pt = total == 0;
        if (fixed_pt) {
            t1 = D.maxt;
        } else {
            if ((k & 1ull) == 0ull) mbg_philox((u32)(k >> 1), 1u, (u32)id, D.variant_id, key0, key1, d4);
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
                th = L.entropy();
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
                if (cur_w >= 0) {
                    flush_cell(D, cur_w, acc_s, acc_p);
                    if (acc_th) atomicAdd(D.th + cur_w, acc_th);
                }
                cur_w = w;
                acc_p = 0;
                acc_th = 0;
                acc_s = s;
            } else if (s != acc_s) {
                flush_cell(D, cur_w, acc_s, acc_p);
                acc_p = 0;
                acc_s = s;
            }
            acc_p += 1;
            acc_th += (u64)rint(ldexp(th_use, D.FT));
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
                    if (cur_w >= 0) {
                        flush_cell(D, cur_w, acc_s, acc_p);
                        if (acc_th) atomicAdd(D.th + cur_w, acc_th);
                    }
                    cur_w = ww;
                    acc_p = 0;
                    acc_th = 0;
                    acc_s = s;
                } else if (s != acc_s) {
                    flush_cell(D, cur_w, acc_s, acc_p);
                    acc_p = 0;
                    acc_s = s;
                }
                acc_p += (u64)rint(ldexp(sl, D.FP));
                acc_th += (u64)rint(ldexp(th_use * sl, D.FT));
                w++;
                b = e;
            }
        }
#endif
        t = t1;

        if (t >= D.maxt) {
            // ---- trajectory end: flush, final state, fixed point, summary ----
            if (cur_w >= 0) {
                flush_cell(D, cur_w, acc_s, acc_p);
                if (acc_th) atomicAdd(D.th + cur_w, acc_th);
            }
            atomicAdd(D.fin + s, 1ull);
            atomicAdd(D.flip_total, flips);
            if (fixed_pt) {
                const u64 slot = atomicAdd(D.fp_count, 1ull);
                if (slot < D.fp_cap)
                    for (int w = 0; w < MBG_NW; w++) D.fp_states[slot * MBG_NW + w] = LANE_X(w);
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

        // ---- flip selection: word sums, then the active bits of one word ----
        const u64 target = mbg_flip_target(dr2, total);
        u64 acc = 0;
        int wsel = MBG_NW - 1;
        bool found = false;
#pragma unroll
        for (int w = 0; w < MBG_NW; w++) {
            const u64 sw = LANE_S(w);
            if (!found) {
                if (acc + sw > target) {
                    wsel = w;
                    found = true;
                } else {
                    acc += sw;
                }
            }
        }
        u64 rem = target - acc;
        u64 m = LANE_A(wsel);
        const u64 xw = LANE_X(wsel);
        const u64 gm = D.gen_mask[wsel];
        // guard for an inconsistent word (never taken when the sums are exact): last active node
        u32 j = m ? (u32)(wsel * 64 + 63 - __clzll((long long)m)) : (u32)(wsel * 64);
        while (m) {
            const u32 b = (u32)__builtin_ctzll(m);
            const u32 cand = (u32)wsel * 64u + b;
            u64 r;
            if ((gm >> b) & 1ull) {
                const mbg_node nd = D.nodes[cand];
                r = L.rate_of(nd, cand, L.key_of(nd, cand));
            } else {
                r = D.rates[2 * cand + ((xw >> b) & 1ull)];
            }
            if (rem < r) {
                j = cand;
                break;
            }
            rem -= r;
            m &= m - 1ull;
        }

        // ---- flip j, re-evaluate its dependents, update sums exactly ----
        LANE_X(j >> 6) = LANE_X(j >> 6) ^ (1ull << (j & 63));
        const mbg_node nj = D.nodes[j];
        for (u32 q = 0; q < nj.dep_cnt; q++) {
            const u32 dp = D.deps[nj.dep_off + q];
            const u32 kk = dp & 0xffffffu;
            const u32 pos = dp >> 24;
            const mbg_node nk = D.nodes[kk];
            const u32 key = L.key_of(nk, kk);
            const u64 rn = L.rate_of(nk, kk, key);
            const u64 ro = L.rate_of(nk, kk, key ^ (1u << pos));
            if (rn != ro) {
                const u32 wk = kk >> 6;
                const u64 bit = 1ull << (kk & 63);
                LANE_S(wk) = LANE_S(wk) + (rn - ro);
                total += rn - ro;
                const u64 a = LANE_A(wk);
                LANE_A(wk) = rn ? (a | bit) : (a & ~bit);
                if (nk.flags & 0x7f) dirty = true;
            }
        }
        if (nj.flags & 0x7f) s ^= 1u << ((nj.flags & 0x7f) - 1);
        flips++;
#if MBG_TRACE
        hash = (hash ^ (u64)j) * 0x100000001b3ull;
#endif
        k++;
    }
}
