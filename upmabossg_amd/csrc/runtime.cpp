// libmbg.so: C ABI (include/mbg.h) over the front end, the model compiler, hipRTC
// code generation and the persistent trajectory kernel.
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <algorithm>
#include <list>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <map>
#include <set>
#include <memory>
#include <numeric>
#include <sstream>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "../../include/mbg.h"
#include "front.h"
#include "model.h"
#include "output.h"
#include "pack.h"
#include "variants.h"

namespace mbg {
template <>
uint64_t* PinnedAllocator<uint64_t>::allocate(size_t n) {
    void* p = nullptr;
    if (hipHostMalloc(&p, std::max<size_t>(1, n) * sizeof(uint64_t), hipHostMallocDefault) != hipSuccess)
        throw std::bad_alloc();
    return static_cast<uint64_t*>(p);
}
template <>
void PinnedAllocator<uint64_t>::deallocate(uint64_t* p, size_t) noexcept {
    (void)hipHostFree(p);
}
}  // namespace mbg

extern const char* const mbg_embedded_math;
extern const char* const mbg_embedded_kernel;

struct mbg_model {
    std::shared_ptr<mbg::ParsedModel> parsed;  // shared by clones
    mbg::CompiledModel cm;
    std::vector<float> mult_up, mult_down, istate;  // empty: model's own values
};

namespace {

using mbg::Blob;
using mbg::compact_fits;
using mbg::Layout;
using mbg::pack_blob;
using mbg::rates_fit_i31;
using mbg::rates_fit_u32;
using mbg::run_scales;
using mbg::RunScales;
using mbg::stage_row;

struct Error : std::runtime_error {
    mbg_status code;
    Error(mbg_status c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw Error(MBG_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

// Timing events, destroyed on every exit path.
struct EventPair {
    hipEvent_t a = nullptr, b = nullptr;
    EventPair() {
        hip_check(hipEventCreate(&a), "hipEventCreate");
        if (hipEventCreate(&b) != hipSuccess) {
            (void)hipEventDestroy(a);
            throw Error(MBG_ERR_HIP, "hipEventCreate failed");
        }
    }
    ~EventPair() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
    EventPair(const EventPair&) = delete;
    EventPair& operator=(const EventPair&) = delete;
};

// A stream capture and what it produces: an error thrown while capturing ends the capture
// (so the stream is usable again) and every event, graph and executable graph is released.
struct CaptureScope {
    hipStream_t stream = nullptr;
    bool capturing = false;
    std::vector<hipEvent_t> events;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~CaptureScope() {
        if (capturing) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(stream, &g);
            if (g) (void)hipGraphDestroy(g);
        }
        for (hipEvent_t e : events)
            if (e) (void)hipEventDestroy(e);
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
};

void copy_err(char* err, size_t n, const std::string& s) {
    if (err && n) {
        std::strncpy(err, s.c_str(), n - 1);
        err[n - 1] = 0;
    }
}


// Kernel argument block; layout identical to struct mbg_desc in device/mbg_kernel.hip.
struct Desc {
    const void* blob;
    const void* rtab;
    const void* istate;
    void* stage;
    void* prob;
    void* th;
    void* fin;
    void* fp_states;
    void* fp_count;
    void* summ;
    void* work_counter;
    void* flip_total;
    uint64_t fp_cap, summ_cap, nsteps, batch, batch_first, first_id, seed;
    double maxt, tick, inv_tick, rscale;
    Layout lay;
    uint32_t variant_id;
    int N, K, W, FP, FT, sigma;
    int compat, pad_;
    void* cells;       // sparse window table (more than 20 readouts): keys, values, misses
    void* cell_vals;
    void* cell_miss;
    uint64_t cell_mask;
    void* lane_state;  // lanes' state in HBM (MBG_STATE_HBM)
    uint32_t lane_stride, pad2_;
    // error columns (MBG_ERRORS): sums of squares, the lanes' window lists, their overflow count
    void* prob2;
    void* th2;
    void* cell_vals2;
    void* err_s;
    void* err_v;
    void* err_miss;
    double sq_scale, sq_scale_th;
    int FQ, FQT;
    uint32_t err_cap, err_stride;
};

using mbg_desc_ptr = const void*;

struct Shape {
    int nw, block, discrete, trace, generic, tables_lds, max_in, dep_batch = 1, scan = 4, wsum_reg = 0, rate32 = 0,
        multi = 0, sumw = 1, compact = 0, wsmad = 0, sparse = 0, prog = 0, th_lds = 0, cells_lds = 0, dep_cap = 0,
        state_hbm = 0, errors = 0, sweep_xcd = 0;
    // MBG_GENCODE: the fast nodes' logic as code generated for this model (appended to the
    // kernel source; empty: the truth tables)
    std::string gen;
    bool operator<(const Shape& o) const {
        return std::tie(nw, block, discrete, trace, generic, tables_lds, max_in, dep_batch, scan, wsum_reg, rate32, multi,
                        sumw, compact, wsmad, sparse, prog, th_lds, cells_lds, dep_cap, state_hbm, errors, sweep_xcd,
                        gen) <
               std::tie(o.nw, o.block, o.discrete, o.trace, o.generic, o.tables_lds, o.max_in, o.dep_batch, o.scan,
                        o.wsum_reg, o.rate32, o.multi, o.sumw, o.compact, o.wsmad, o.sparse, o.prog, o.th_lds,
                        o.cells_lds, o.dep_cap, o.state_hbm, o.errors, o.sweep_xcd, o.gen);
    }
};

// The logic of every node as generated code, the reference's way of compiling a model
// (src/generator.cpp:84-122): one case per node in a switch on the node id, its `logic`
// expression written out in the FP32 arithmetic the truth tables are built with
// (model.cpp, StateEval).  Empty when some node has no plain logic attribute (generic rates,
// logic programs) or an alias recursion.
struct GenLogic {
    const mbg::ParsedModel& pm;
    int self = 0, depth = 0;
    bool ok = true;

    static std::string lit(float v) {
        char b[64];
        std::snprintf(b, sizeof b, "%.9gf", (double)v);
        std::string t(b);
        if (t == "inff") return "__builtin_inff()";
        if (t == "-inff") return "(-__builtin_inff())";
        if (t == "nanf" || t == "-nanf") return "__builtin_nanf(\"\")";
        return "(" + t + ")";
    }
    std::string emit(const mbg::Ast* e) {
        using mbg::Ast;
        if (!ok) return "0.f";
        switch (e->kind) {
            case Ast::Number: return lit(e->number);
            case Ast::Variable: return lit(mbg::evaluate(e, pm));
            case Ast::Group: return emit(e->kids[0].get());
            case Ast::Node: {
                const int i = pm.index_of(e->text);
                if (i < 0) ok = false;
                return "(float)L.xbit(" + std::to_string(i < 0 ? 0 : i) + "u)";
            }
            case Ast::Alias: {
                const mbg::Ast* a = pm.nodes[self].attr(e->text.substr(1));
                if (!a || ++depth > 64) {
                    ok = false;
                    return "0.f";
                }
                const std::string v = emit(a);
                depth--;
                return v;
            }
            case Ast::Unary: {
                const std::string v = emit(e->kids[0].get());
                if (e->oper == Ast::Sub) return "(-" + v + ")";
                if (e->oper == Ast::Not) return "(" + v + " == 0.f ? 1.f : 0.f)";
                return v;
            }
            case Ast::Ternary:
                return "(" + emit(e->kids[0].get()) + " != 0.f ? " + emit(e->kids[1].get()) + " : " +
                       emit(e->kids[2].get()) + ")";
            case Ast::Binary: {
                const std::string l = emit(e->kids[0].get()), r = emit(e->kids[1].get());
                switch (e->oper) {
                    case Ast::Add: return "(" + l + " + " + r + ")";
                    case Ast::Sub: return "(" + l + " - " + r + ")";
                    case Ast::Mul: return "(" + l + " * " + r + ")";
                    case Ast::Div: return "(" + l + " / " + r + ")";
                    case Ast::And: return "((" + l + " != 0.f && " + r + " != 0.f) ? 1.f : 0.f)";
                    case Ast::Or: return "((" + l + " != 0.f || " + r + " != 0.f) ? 1.f : 0.f)";
                    case Ast::Xor: return "(((" + l + " != 0.f) != (" + r + " != 0.f)) ? 1.f : 0.f)";
                    case Ast::Eq: return "(" + l + " == " + r + " ? 1.f : 0.f)";
                    case Ast::Ne: return "(" + l + " != " + r + " ? 1.f : 0.f)";
                    case Ast::Lt: return "(" + l + " < " + r + " ? 1.f : 0.f)";
                    case Ast::Le: return "(" + l + " <= " + r + " ? 1.f : 0.f)";
                    case Ast::Gt: return "(" + l + " > " + r + " ? 1.f : 0.f)";
                    case Ast::Ge: return "(" + l + " >= " + r + " ? 1.f : 0.f)";
                    default: ok = false; return "0.f";
                }
            }
        }
        ok = false;
        return "0.f";
    }
};

std::string generate_logic_source(const mbg::ParsedModel& pm, const mbg::CompiledModel& cm) {
    if (cm.any_generic || cm.any_prog) return std::string();
    GenLogic g{pm};
    std::ostringstream os;
    os << "__device__ u32 mbg_gen_logic(u32 i, const Lane& L) {\n    switch (i) {\n";
    for (int i = 0; i < cm.N; i++) {
        const mbg::Ast* e = pm.nodes[i].attr("logic");
        if (!e) return std::string();
        g.self = i;
        g.depth = 0;
        const std::string v = g.emit(e);
        if (!g.ok) return std::string();
        os << "    case " << i << "u: return " << v << " != 0.f ? 1u : 0u;\n";
    }
    os << "    default: return 0u;\n    }\n}\n";
    return os.str();
}

std::string generate_source(const mbg::CompiledModel& cm, const Shape& sh) {
    std::ostringstream os;
    os << "// Trajectory kernels generated for a model of " << cm.N << " nodes (" << cm.K
       << " non-internal), " << (cm.discrete ? "discrete" : "continuous") << " time.\n";
    os << "#define MBG_NW " << sh.nw << "\n";
    os << "#define MBG_BLOCK " << sh.block << "\n";
    os << "#define MBG_DISCRETE " << sh.discrete << "\n";
    os << "#define MBG_TRACE " << sh.trace << "\n";
    os << "#define MBG_GENERIC " << sh.generic << "\n";
    os << "#define MBG_TABLES_LDS " << sh.tables_lds << "\n";
    os << "#define MBG_MAX_IN " << sh.max_in << "\n";
    os << "#define MBG_DEP_BATCH " << sh.dep_batch << "\n";
    os << "#define MBG_SCAN " << sh.scan << "\n";
    os << "#define MBG_WSUM_REG " << sh.wsum_reg << "\n";
    os << "#define MBG_RATE32 " << sh.rate32 << "\n";
    os << "#define MBG_MULTI " << sh.multi << "\n";
    os << "#define MBG_SUMW " << sh.sumw << "\n";
    os << "#define MBG_COMPACT " << sh.compact << "\n";
    os << "#define MBG_WSMAD " << sh.wsmad << "\n";
    os << "#define MBG_SPARSE " << sh.sparse << "\n";
    os << "#define MBG_PROG " << sh.prog << "\n";
    os << "#define MBG_TH_LDS " << sh.th_lds << "\n";
    os << "#define MBG_CELLS_LDS " << sh.cells_lds << "\n";
    os << "#define MBG_DEP_CAP " << sh.dep_cap << "\n";
    os << "#define MBG_STATE_HBM " << sh.state_hbm << "\n";
    os << "#define MBG_ERRORS " << sh.errors << "\n";
    os << "#define MBG_SWEEP_XCD " << sh.sweep_xcd << "\n";
    os << "#define MBG_GENCODE " << (sh.gen.empty() ? 0 : 1) << "\n";
    os << "#define MBG_HD __device__ __forceinline__\n";
    os << mbg_embedded_math << "\n" << mbg_embedded_kernel << "\n" << sh.gen;
    return os.str();
}

struct Module {
    hipModule_t mod = nullptr;
    hipFunction_t sim = nullptr, init_state = nullptr, init_rates = nullptr, init_readouts = nullptr,
                  marginals = nullptr, marginals_batch = nullptr, init_rates_multi = nullptr, fp_dedup = nullptr,
                  fp_compact = nullptr,
                  fp_gather = nullptr, cells_count = nullptr, cells_compact = nullptr;
};

// Device buffer kept by the context across runs; it only grows.
template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t cap = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    void alloc(size_t count) {
        count = std::max<size_t>(1, count);
        if (p && count <= cap) return;
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        hip_check(hipMalloc((void**)&p, count * sizeof(T)), "hipMalloc");
        cap = count;
    }
    void upload(const std::vector<T>& v, hipStream_t s) {
        alloc(v.size());
        if (!v.empty()) hip_check(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s), "hipMemcpyAsync");
    }
    void zero(size_t count, hipStream_t s) { fill(count, 0, s); }
    void fill(size_t count, int byte, hipStream_t s) {
        alloc(count);
        hip_check(hipMemsetAsync(p, byte, std::max<size_t>(1, count) * sizeof(T), s), "hipMemsetAsync");
    }
};

struct RunBuffers {
    DevBuf<uint64_t> blob, rtab, stage, prob, th, fin, fp, fpc, summ, work, flips, cells, cell_vals, cell_miss, cell_out,
        lane_state, prob2, th2, cell_vals2, err_s, err_v, err_miss;
    DevBuf<float> istate;
};

// One variant of a sweep: its tables, device buffers and kernel descriptor.
struct VariantRun {
    const mbg_model* m = nullptr;
    int NWc = 0;        // words of its class (state rows are NWc words wide)
    uint64_t n = 0;
    RunBuffers buf;
    Desc D{};
    mbg_run_info info{};
    std::vector<uint64_t> fixed_states, fixed_counts;
    std::vector<uint64_t> marg_raw;  // [W][K] readout marginals (exact sums), from the sweep graph
    const Module* mod = nullptr;     // its class's module (fixed-point deduplication)
};

// Host mirror of struct mbg_marg_job in device/mbg_kernel.hip.
struct MargJob {
    const uint64_t* prob;
    uint64_t out_off;
    int W, K;
};

// The measured layout decisions of one sweep class (-1: the default rule): tables staged per
// block in LDS or read from HBM, the active sets in HBM (2) or LDS (0), dependents re-evaluated
// per loop iteration (0: no cap).
struct ClassChoice {
    int tables_lds = -1, state_hbm = -1, dep_cap = -1;
};

struct SweepClass {
    Shape sh{};
    std::vector<long long> key;  // word class, time mode, error columns, structure of the members
    hipFunction_t sim = nullptr, init_state = nullptr, init_rates = nullptr, init_readouts = nullptr,
                  marginals = nullptr, marginals_batch = nullptr, init_rates_multi = nullptr;
    std::vector<int> members;
    DevBuf<Desc> descs;
    size_t smem = 0;
    int blocks_per_cu = 1;
    int wmax = 1;  // largest window count of the class's variants (per-wave entropy partials)
    DevBuf<uint64_t> act;  // the lanes' active sets (MBG_STATE_HBM 2) or whole state (1) in HBM
    DevBuf<uint64_t> err_s, err_v, err_miss;  // error columns: the lanes' window lists, overflows
    double work = 0;         // launch-order estimate: trajectories x nodes x max_time over the members
    double kernel_ms = 0;    // the class kernel's span in the last graph (first block start to last end)
    uint64_t trajectories = 0, transitions = 0;
    bool tuned = false;      // the shape came from mbg_autotune_variants
    uint32_t tab_bytes = 0;  // tables in LDS: the class's largest table blob, staged per block
    uint64_t grid = 1;       // blocks of the class kernel
    DevBuf<int> block_start; // tables in LDS: each block's first variant (index into the members)
};

struct SweepState {
    std::vector<std::unique_ptr<VariantRun>> vars;
    std::vector<std::unique_ptr<SweepClass>> classes;
    mbg_sweep_info info{};
};

// Device buffers shared by all variants of a sweep, kept by the context across sweeps.
struct SweepShared {
    DevBuf<uint64_t> counters;  // [variant][2]: fixed-point count, transitions
    DevBuf<MargJob> jobs;
    DevBuf<uint64_t> marg;      // all variants' marginals, [variant][W][K]
    DevBuf<uint64_t> clk;       // [class][2]: first block start, last block end (device clock)
};

}  // namespace

struct mbg_ctx {
    int device = 0;
    int cus = 0;
    int lds_per_block = 65536;  // device limit on (static + dynamic) LDS per workgroup
    int wall_khz = 0;           // rate of the device's constant clock (wall_clock64)
    hipStream_t stream = nullptr;
    std::string err;
    std::map<Shape, Module> modules;
    std::map<std::vector<long long>, Shape> tuned;  // measured best shape per model structure
    std::map<std::vector<long long>, ClassChoice> tuned_classes;  // measured choices per sweep class
    const ClassChoice* force_class = nullptr;  // the sweep autotuner's candidate (every class)
    // transitions per trajectory measured by earlier sweeps, per variant (model, trajectories,
    // seed, variant id): the work estimate that orders a class's variants and shares its blocks
    std::map<std::vector<uint64_t>, double> variant_cost;
    RunBuffers buf;
    std::unique_ptr<SweepState> sweep;       // last variant sweep
    std::vector<std::unique_ptr<VariantRun>> variant_pool;  // device buffers reused by later sweeps
    SweepShared sweep_shared;
    // fixed-point deduplication on the device
    DevBuf<uint32_t> fp_keys, fp_idx;
    DevBuf<uint64_t> fp_counts, fp_cnt, fp_n, fp_out;
    mbg::PinnedU64 fp_host, fp_host_cnt;  // distinct fixed points read back (kept: no refaults)
    std::vector<hipStream_t> class_streams;  // one per word class of a sweep
    // last run
    bool have = false;
    mbg_run_info info{};
    mbg::PinnedU64 prob;  // window table, pinned for a fast device-to-host copy
    bool prob_on_host = false;  // the window table stays on the device until a getter asks
    hipFunction_t marg_fn = nullptr;  // mbg_marginals of the last run's module
    DevBuf<uint64_t> marg;
    std::vector<uint64_t> th, fin, fixed_states, fixed_counts, summ;
    // more than 20 readouts: non-zero window cells (window-major, states ascending) and final
    // states, read back from the device hash table; cell_cap is the table size to start from
    bool sparse = false;
    std::vector<uint32_t> cell_w;
    std::vector<uint64_t> cell_s, cell_v, fin_s, fin_v;
    uint64_t cell_cap = 0;
    // error columns of the last run (when asked for): sums of squares, window-major; the
    // sparse form follows cell_v
    std::vector<uint64_t> prob2, th2, cell_v2;
    uint32_t err_cap = 8;  // entries of each lane's window list to start the next run with
};

namespace {

const Module& get_module(mbg_ctx* c, const mbg::CompiledModel& cm, const Shape& sh) {
    auto it = c->modules.find(sh);
    if (it != c->modules.end()) return it->second;
    const std::string src = generate_source(cm, sh);
    hiprtcProgram prog;
    if (hiprtcCreateProgram(&prog, src.c_str(), "mbg_kernel.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS)
        throw Error(MBG_ERR_COMPILE, "hiprtcCreateProgram failed");
    // -O1: with -O2 and -O3 some shapes of these kernels (a 177-node model in blocks of 512 or
    // 1024 threads, the 4-, 24- and 28-word sweep classes) gave wrong transition counts that
    // changed from run to run on gfx950, while the kernel source runs bit-exact in the host
    // emulation (also under AddressSanitizer and MemorySanitizer); -O1 code of the same shapes
    // is bit-exact and deterministic
    std::vector<const char*> opts = {"--offload-arch=gfx950", "-O1", "-std=c++17", "-ffp-contract=off",
                                     "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
    // MBG_RTC_OPTS: extra compiler options, space-separated (code-generation experiments)
    std::vector<std::string> extra;
    if (const char* e = std::getenv("MBG_RTC_OPTS")) {
        std::istringstream is(e);
        for (std::string o; is >> o;) extra.push_back(o);
    }
    for (const std::string& o : extra) opts.push_back(o.c_str());
    const hiprtcResult rc = hiprtcCompileProgram(prog, (int)opts.size(), opts.data());
    size_t log_size = 0;
    hiprtcGetProgramLogSize(prog, &log_size);
    std::string log(log_size, '\0');
    if (log_size) hiprtcGetProgramLog(prog, &log[0]);
    if (rc != HIPRTC_SUCCESS) {
        hiprtcDestroyProgram(&prog);
        throw Error(MBG_ERR_COMPILE, "hipRTC compilation failed: " + log);
    }
    size_t code_size = 0;
    hiprtcGetCodeSize(prog, &code_size);
    std::vector<char> code(code_size);
    hiprtcGetCode(prog, code.data());
    hiprtcDestroyProgram(&prog);
    Module m;
    hip_check(hipModuleLoadData(&m.mod, code.data()), "hipModuleLoadData");
    hip_check(hipModuleGetFunction(&m.sim, m.mod, sh.multi ? "mbg_simulate_multi" : "mbg_simulate"),
              "hipModuleGetFunction");
    hip_check(hipModuleGetFunction(&m.marginals, m.mod, "mbg_marginals"), "hipModuleGetFunction");
    hip_check(hipModuleGetFunction(&m.marginals_batch, m.mod, "mbg_marginals_batch"), "hipModuleGetFunction");
    hip_check(hipModuleGetFunction(&m.fp_dedup, m.mod, "mbg_fp_dedup"), "hipModuleGetFunction");
    hip_check(hipModuleGetFunction(&m.fp_compact, m.mod, "mbg_fp_compact"), "hipModuleGetFunction");
    hip_check(hipModuleGetFunction(&m.fp_gather, m.mod, "mbg_fp_gather"), "hipModuleGetFunction");
    hip_check(hipModuleGetFunction(&m.cells_count, m.mod, "mbg_cells_count"), "hipModuleGetFunction");
    hip_check(hipModuleGetFunction(&m.cells_compact, m.mod, "mbg_cells_compact"), "hipModuleGetFunction");
    if (sh.multi && sh.nw <= 128)
        hip_check(hipModuleGetFunction(&m.init_rates_multi, m.mod, "mbg_init_rates_multi"), "hipModuleGetFunction");
    hip_check(hipModuleGetFunction(&m.init_state, m.mod, "mbg_init_state"), "hipModuleGetFunction");
    hip_check(hipModuleGetFunction(&m.init_rates, m.mod, "mbg_init_rates"), "hipModuleGetFunction");
    hip_check(hipModuleGetFunction(&m.init_readouts, m.mod, "mbg_init_readouts"), "hipModuleGetFunction");
    return c->modules.emplace(sh, m).first->second;
}


// LDS of the lanes' state: x and act, plus the word rate sums unless they are in VGPRs only;
// x alone with act (state_hbm 2) or act and the word sums (3) in HBM, nothing with the whole
// state in HBM (state_hbm 1)
size_t lane_lds_bytes(int nw, int block, int wsum_reg, int state_hbm = 0) {
    if (state_hbm == 1) return 0;
    return (size_t)(state_hbm >= 2 ? 1 : wsum_reg == 1 ? 2 : 3) * nw * block * 8;
}

// 64-bit words per lane of the HBM lane-state buffer
int lane_hbm_words(int nw, int wsum_reg, int state_hbm) {
    return (state_hbm == 2 ? 1 : state_hbm == 3 ? 2 : wsum_reg == 1 ? 2 : 3) * nw;
}

int env_int(const char* name, int dflt) {
    const char* e = std::getenv(name);
    return e ? std::atoi(e) : dflt;
}

// Word rate sums live in VGPRs (an unrolled select over at most 16 sums per update) when
// the model has at most 16 words, or up to 32 words in groups of two (an even word count).
// Otherwise they take a third LDS array beside x and act, summed again in at most 16 group
// sums in VGPRs (mode 2) so that a selection does not scan every word.
bool reg_sums_fit(int nw) { return nw <= 16 || (nw <= 32 && nw % 2 == 0); }
// 1: word sums in VGPRs; 2 (wider models): word sums in LDS under at most 16 group sums in
// VGPRs, so a selection scans the groups and then one group's words; 0: LDS only.
int default_wsum_reg(int nw) {
    int v = env_int("MBG_TUNE_WSUM_REG", reg_sums_fit(nw) ? 1 : 2);
    if (v == 1 && !reg_sums_fit(nw)) v = 2;
    return v < 0 || v > 2 ? 0 : v;
}

// Words per register rate-sum group: fewer, wider groups make each sum update cheaper (an
// unrolled select over the groups) at the cost of scanning more act words per selection.
int default_sumw(int nw, int wsum_reg) {
    if (!wsum_reg) return 1;
    if (wsum_reg == 2) return (nw + 15) / 16;  // at most 16 groups, the last one possibly short
    int sw = env_int("MBG_TUNE_SUMW", nw > 16 ? 2 : 1);
    const int max_groups = env_int("MBG_TUNE_MAX_GROUPS", 16);
    if (sw < 1) sw = 1;
    while (sw > 1 && nw % sw) sw /= 2;
    while (nw / sw > max_groups && nw % (sw * 2) == 0) sw *= 2;
    return sw;
}

// Dependents re-evaluated per loop iteration (0: all of them).  A wave runs the dependents
// loop as many rounds as its lane with the most dependents needs; with a cap, such a lane
// finishes over the next iterations (and steps later) while the others go on stepping.
// The cap pays where a few nodes have many more dependents than the mean (random networks:
// roughly Poisson out-degrees); where every node has about as many (the ring models), it only
// splits each flip's work over more iterations.  Default: 1.5 times the mean dependent count,
// at least 4 and at least two batches; MBG_TUNE_DEP_CAP sets it (0 removes the cap).  The last
// round of a capped iteration may be a partial batch.
// Dependents re-evaluated per round of independent table reads: always 1.  With 2 or more
// (MBG_DEP_BATCH > 1) the GPU gave wrong transition counts that changed from run to run in
// several launch shapes (with kernels compiled at -O3: a 177-node model in blocks of 512 and
// 1,024, 4-, 20-, 24- and 28-word sweep classes, 60- and 130-node models; at -O1: a 4-word sweep
// class with its tables per block in LDS and blocks of 1,024), while the same kernel source runs
// bit-exact in the CPU emulation of tests/kernel_emu for those shapes and batches.  Every case
// that went wrong had batches of 2; with 1 every case is exact.  The cause is not identified, so
// the layout is refused (MBG_TUNE_DEP_BATCH may only ask for 1).
int dependents_per_round() {
    const int v = env_int("MBG_TUNE_DEP_BATCH", 1);
    if (v != 1) throw Error(MBG_ERR_LIMIT, "MBG_TUNE_DEP_BATCH: only one dependent per round is supported");
    return 1;
}

int default_dep_cap(const mbg::CompiledModel& cm, int dep_batch) {
    double sum = 0;
    for (const mbg::NodeRec& n : cm.nodes) sum += n.dep_cnt;
    const double mean = cm.N ? sum / cm.N : 0.0;
    const int dflt = std::max({4, 2 * dep_batch, (int)std::ceil(1.5 * mean)});
    const int v = env_int("MBG_TUNE_DEP_CAP", dflt);
    return v <= 0 ? 0 : v;
}

// Launch shape: block size, table placement (LDS or HBM) and word-sum placement (LDS or
// VGPRs).  Without tuning knobs: stage the tables in LDS when they fit beside the lanes'
// state and keep the most waves resident per CU (160 KiB of LDS per CU on gfx950).
Shape choose_shape(const mbg_ctx* c, const mbg::CompiledModel& cm, uint32_t blob_bytes, int block_req, int trace,
                   int rate32, bool lanes_in_lds = true) {
    const size_t cu_lds = 160 * 1024;
    const int max_in = std::max(1, cm.max_inputs);
    // word sums in VGPRs pay off while the unrolled select over the words stays short
    const int wsum_reg = default_wsum_reg(cm.NW);
    const int tables_env = env_int("MBG_TUNE_TABLES", -1);
    if (!block_req) block_req = env_int("MBG_TUNE_BLOCK", 0);
    Shape best{cm.NW, 64, cm.discrete ? 1 : 0, trace ? 1 : 0, cm.any_generic ? 1 : 0, 0, max_in};
    best.wsum_reg = wsum_reg;
    best.rate32 = rate32;
    int best_score = -1;
    for (int tables = 1; tables >= 0; tables--) {
        if (tables_env >= 0 && tables != tables_env) continue;
        for (int block = 1024; block >= 64; block -= 64) {
            if (block_req && block != block_req) continue;
            const size_t lds = lane_lds_bytes(cm.NW, block, wsum_reg) + (tables ? blob_bytes : 0);
            if (lds > (size_t)c->lds_per_block) continue;
            const int waves = std::min(16, (int)(cu_lds / lds) * (block / 64));
            const int score = tables ? waves * 2 : waves;  // tables in LDS are worth a modest occupancy loss
            if (waves > 0 && score > best_score) {
                best_score = score;
                best.block = block;
                best.tables_lds = tables;
            }
        }
    }
    best.dep_batch = dependents_per_round();
    best.dep_cap = default_dep_cap(cm, best.dep_batch);
    best.scan = env_int("MBG_TUNE_SCAN", 4) >= 8 ? 8 : 4;
    best.sumw = default_sumw(cm.NW, best.wsum_reg);
    if (best_score < 0 && lanes_in_lds) {
        if (block_req) throw Error(MBG_ERR_LIMIT, "the requested block does not fit in LDS for this model");
        throw Error(MBG_ERR_LIMIT, "model too large: per-lane state does not fit in LDS");
    }
    return best;
}

// The block with the most resident waves for a layout (word-sum placement, tables in LDS or
// not), optionally only among blocks with fewer than below_waves waves; 0 if nothing fits.
int layout_block(const mbg_ctx* c, int nw, int wsum_reg, uint32_t blob_bytes, int tables, int* waves_out,
                 int below_waves = 1 << 30, int state_hbm = 0) {
    const size_t cu_lds = 160 * 1024;
    int best = 0, best_waves = 0;
    for (int block = 1024; block >= 64; block -= 64) {
        const size_t lds = lane_lds_bytes(nw, block, wsum_reg, state_hbm) + (tables ? blob_bytes : 0);
        if (lds > (size_t)c->lds_per_block) continue;
        const int waves = std::min(16, (int)(cu_lds / lds) * (block / 64));
        if (waves >= below_waves) continue;
        if (waves > best_waves) {
            best_waves = waves;
            best = block;
        }
    }
    if (waves_out) *waves_out = best_waves;
    return best;
}

// Block for the active sets in HBM (MBG_STATE_HBM 2): x in LDS beside fixed_bytes per block
// and per_wave_bytes per wave.  At most 12 waves per block: launch bounds of 768 threads leave
// the step loop about 170 VGPRs (3 waves per SIMD), while 16 would cap it at 128 and spill, so
// a CU holds at most 12 waves of this kernel whatever the LDS allows.  0 if nothing fits.
// The layout needs the rate sums in registers (x alone stays in LDS): one sum per word up to
// 16 words, or per group of two words up to 32.  (With groups of two, a sweep of the 20-word
// class once miscounted with kernels compiled at -O3; at -O1, see get_module, that sweep and
// single runs of 1,300-2,000-node models in this layout are bit-exact.)
bool act_hbm_allowed(int wsum_reg, int sumw) {
    (void)sumw;
    return wsum_reg == 1;
}

int act_hbm_block(const mbg_ctx* c, int nw, size_t fixed_bytes, size_t per_wave_bytes, int* waves_out) {
    const size_t cu_lds = 160 * 1024;
    int best = 0, best_waves = 0;
    for (int block = 768; block >= 64; block -= 64) {
        const size_t lds = lane_lds_bytes(nw, block, 1, 2) + fixed_bytes + (size_t)(block / 64) * per_wave_bytes;
        if (lds > (size_t)c->lds_per_block) continue;
        const int waves = std::min(12, (int)(cu_lds / lds) * (block / 64));
        if (waves > best_waves) {
            best_waves = waves;
            best = block;
        }
    }
    if (waves_out) *waves_out = best_waves;
    return best;
}

void launch(hipFunction_t f, uint64_t grid, int block, size_t smem, hipStream_t s, Desc& D, const char* what) {
    void* args[] = {&D};
    hip_check(hipModuleLaunchKernel(f, (unsigned)grid, 1, 1, (unsigned)block, 1, 1, (unsigned)smem, s, args, nullptr), what);
}

// Every device buffer a descriptor names must be allocated before a launch reads it.
void check_desc(const Desc& D, bool sparse, bool state_hbm = false) {
    const void* need[] = {D.blob, D.istate, D.stage, D.prob, D.th, D.fin, D.fp_states, D.fp_count, D.summ,
                          D.work_counter, D.flip_total};
    for (const void* q : need)
        if (!q) throw Error(MBG_ERR_ARG, "internal error: a device buffer of the run is not allocated");
    if (sparse && (!D.cells || !D.cell_vals || !D.cell_miss))
        throw Error(MBG_ERR_ARG, "internal error: the sparse window table is not allocated");
    if (state_hbm && (!D.lane_state || !D.lane_stride))
        throw Error(MBG_ERR_ARG, "internal error: the lanes' state buffer is not allocated");
}


// Distinct fixed points in the reference order (src/state.h:43-49): rows of `words` words,
// of which the first `keep` are returned per state.
void unique_fixed_points(const std::vector<uint64_t>& fps, int words, int keep, std::vector<uint64_t>& states,
                         std::vector<uint64_t>& counts) {
    const uint64_t nfp = fps.size() / (size_t)words;
    std::vector<uint64_t> idx(nfp);
    std::iota(idx.begin(), idx.end(), 0);
    auto row = [&](uint64_t r) { return fps.data() + r * words; };
    std::sort(idx.begin(), idx.end(), [&](uint64_t a, uint64_t b) {
        const uint64_t *x = row(a), *y = row(b);
        for (int w = words - 1; w >= 0; w--)
            if (x[w] != y[w]) return x[w] < y[w];
        return false;
    });
    states.clear();
    counts.clear();
    for (size_t i = 0; i < idx.size();) {
        size_t j = i;
        while (j < idx.size() && std::equal(row(idx[i]), row(idx[i]) + words, row(idx[j]))) j++;
        states.insert(states.end(), row(idx[i]), row(idx[i]) + keep);
        counts.push_back(j - i);
        i = j;
    }
}

// Distinct fixed points in the reference order (most significant word first, src/state.h:43-49)
// from nfp records of `stride` words on the device (the first `keep` significant).  Few records
// are sorted on the host; many are first reduced to distinct states and counts on the device
// (mbg_fp_dedup / _compact / _gather), so only those cross PCIe.
void collect_fixed_points(mbg_ctx* c, const Module& mod, const uint64_t* d_fp, uint64_t nfp, int stride, int keep,
                          hipStream_t s, std::vector<uint64_t>& states, std::vector<uint64_t>& counts) {
    if (nfp == 0) {
        states.clear();
        counts.clear();
        return;
    }
    // MBG_TUNE_FP_HOST=1: every record to the host and sorted there (the check of the device path)
    if (nfp <= 4096 || env_int("MBG_TUNE_FP_HOST", 0) != 0) {
        std::vector<uint64_t> fps((size_t)nfp * stride);
        hip_check(hipMemcpy(fps.data(), d_fp, fps.size() * 8, hipMemcpyDeviceToHost), "hipMemcpy");
        unique_fixed_points(fps, stride, keep, states, counts);
        return;
    }
    uint64_t cap = 1;
    while (cap < 2 * nfp) cap <<= 1;
    c->fp_keys.alloc(cap);
    hip_check(hipMemsetAsync(c->fp_keys.p, 0xFF, cap * 4, s), "hipMemsetAsync");
    c->fp_counts.zero(cap, s);
    c->fp_idx.alloc(nfp);
    c->fp_cnt.alloc(nfp);
    c->fp_n.zero(1, s);
    int st = stride, kw = keep;
    uint64_t n = nfp, mask = cap - 1;
    const void* fp = d_fp;
    void* keys = c->fp_keys.p;
    void* cnts = c->fp_counts.p;
    void* a1[] = {(void*)&fp, &n, &st, &kw, &keys, &cnts, &mask};
    hip_check(hipModuleLaunchKernel(mod.fp_dedup, (unsigned)((nfp + 255) / 256), 1, 1, 256, 1, 1, 0, s, a1, nullptr),
              "launch mbg_fp_dedup");
    void* oidx = c->fp_idx.p;
    void* ocnt = c->fp_cnt.p;
    void* on = c->fp_n.p;
    void* a2[] = {&keys, &cnts, &cap, &oidx, &ocnt, &on};
    hip_check(hipModuleLaunchKernel(mod.fp_compact, (unsigned)((cap + 255) / 256), 1, 1, 256, 1, 1, 0, s, a2, nullptr),
              "launch mbg_fp_compact");
    uint64_t m = 0;
    hip_check(hipMemcpyAsync(&m, c->fp_n.p, 8, hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
    hip_check(hipStreamSynchronize(s), "fixed-point deduplication");
    c->fp_out.alloc(m * (uint64_t)keep);
    void* out = c->fp_out.p;
    void* a3[] = {(void*)&fp, &st, &kw, &oidx, &m, &out};
    hip_check(hipModuleLaunchKernel(mod.fp_gather, (unsigned)((m * keep + 255) / 256), 1, 1, 256, 1, 1, 0, s, a3, nullptr),
              "launch mbg_fp_gather");
    // into pinned buffers kept by the context (a full-speed copy, and no page faults or zero
    // fill on later runs)
    if (c->fp_host.size() < (size_t)m * keep) c->fp_host.resize((size_t)m * keep);
    if (c->fp_host_cnt.size() < m) c->fp_host_cnt.resize(m);
    const uint64_t* dist = c->fp_host.data();
    const uint64_t* cnt = c->fp_host_cnt.data();
    hip_check(hipMemcpyAsync(c->fp_host.data(), c->fp_out.p, (size_t)m * keep * 8, hipMemcpyDeviceToHost, s),
              "hipMemcpyAsync");
    hip_check(hipMemcpyAsync(c->fp_host_cnt.data(), c->fp_cnt.p, m * 8, hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
    hip_check(hipStreamSynchronize(s), "fixed-point gather");
    // ordered by a key of the two most significant words, the whole rows deciding ties; large
    // sets sorted in slices on host threads, then merged
    struct Key {
        uint64_t a, b, r;
    };
    auto row = [&](uint64_t r) { return dist + r * keep; };
    std::vector<Key> ks(m);
    for (uint64_t r = 0; r < m; r++) ks[r] = {row(r)[keep - 1], keep > 1 ? row(r)[keep - 2] : 0ull, r};
    auto less = [&](const Key& x, const Key& y) {
        if (x.a != y.a) return x.a < y.a;
        if (x.b != y.b) return x.b < y.b;
        const uint64_t *p = row(x.r), *q = row(y.r);
        for (int w = keep - 3; w >= 0; w--)
            if (p[w] != q[w]) return p[w] < q[w];
        return false;
    };
    const int nt = m >= 65536 ? (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency())) : 1;
    std::vector<size_t> off(nt + 1);
    for (int t = 0; t <= nt; t++) off[t] = (size_t)(m * (uint64_t)t / (uint64_t)nt);
    auto on_threads = [&](auto&& f) {
        std::vector<std::thread> th;
        for (int t = 1; t < nt; t++) th.emplace_back(f, t);
        f(0);
        for (auto& x : th) x.join();
    };
    on_threads([&](int t) { std::sort(ks.begin() + (long)off[t], ks.begin() + (long)off[t + 1], less); });
    for (int w = 1; w < nt; w *= 2) {
        std::vector<std::thread> th;
        for (int t = 0; t + w < nt; t += 2 * w)
            th.emplace_back([&, t, w] {
                std::inplace_merge(ks.begin() + (long)off[t], ks.begin() + (long)off[t + w],
                                   ks.begin() + (long)off[std::min(t + 2 * w, nt)], less);
            });
        for (auto& x : th) x.join();
    }
    // every element is written below: resizing a vector that held as many before fills nothing
    states.resize((size_t)m * keep);
    counts.resize(m);
    on_threads([&](int t) {
        for (size_t k = off[t]; k < off[t + 1]; k++) {
            std::memcpy(states.data() + k * keep, row(ks[k].r), (size_t)keep * 8);
            counts[k] = cnt[ks[k].r];
        }
    });
}

// Structure of a model that decides which launch shapes are possible and how they perform.
std::vector<long long> tune_key(const mbg::CompiledModel& cm, const mbg::RateTables& rt, int trace) {
    return {cm.NW, cm.N, cm.K, cm.discrete ? 1 : 0, trace ? 1 : 0, cm.any_generic ? 1 : 0, cm.max_inputs,
            compact_fits(cm) ? 1 : 0, rates_fit_u32(rt) ? 1 : 0, rates_fit_i31(rt) ? 1 : 0,
            (long long)cm.deps.size(), (long long)cm.inputs.size(), (long long)cm.lut.size()};
}

const char* const kSparseMessage =
    "the run has more than 20 readouts: its window cells and final states are sparse "
    "(mbg_get_probtraj_cells, mbg_get_final_cells)";

// Largest sparse window table: 2^30 slots (16 GiB of keys and values).
constexpr uint64_t kMaxCells = 1ull << 30;

// Reads the sparse table back: the taken slots are counted and compacted on the device
// (mbg_cells_count, mbg_cells_compact) and copied, then sorted by key (window-major, states ascending) with a counting sort on the
// window followed by per-window sorts on up to 16 host threads; the keys of window W are the
// final states.  The next sparse run starts from twice the size when more than half the slots
// were used (short probe sequences).
void read_cells(mbg_ctx* c, const Module& mod, uint64_t cap, int W, int K, hipStream_t s, bool squares) {
    RunBuffers& B = c->buf;
    const unsigned blocks = (unsigned)((cap + 255) / 256);
    uint64_t n = 0;
    {
        B.cell_out.alloc(1);
        hip_check(hipMemsetAsync(B.cell_out.p, 0, 8, s), "hipMemsetAsync");
        const void* kp = B.cells.p;
        uint64_t capv = cap;
        void* on = B.cell_out.p;
        void* args[] = {(void*)&kp, &capv, &on};
        hip_check(hipModuleLaunchKernel(mod.cells_count, blocks, 1, 1, 256, 1, 1, 0, s, args, nullptr),
                  "launch mbg_cells_count");
        hip_check(hipMemcpyAsync(&n, B.cell_out.p, 8, hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
        hip_check(hipStreamSynchronize(s), "mbg_cells_count");
    }
    if (n > cap) throw Error(MBG_ERR_HIP, "sparse window table: inconsistent slot count");
    std::vector<uint64_t> keys(n), vals(n), vals2(squares ? n : 0);
    if (n) {
        B.cell_out.alloc(3 * n + 1);
        hip_check(hipMemsetAsync(B.cell_out.p + 3 * n, 0, 8, s), "hipMemsetAsync");
        const void* kp = B.cells.p;
        const void* vp = B.cell_vals.p;
        const void* vp2 = squares ? (const void*)B.cell_vals2.p : nullptr;
        void* ok = B.cell_out.p;
        void* ov = B.cell_out.p + n;
        void* ov2 = B.cell_out.p + 2 * n;
        void* on = B.cell_out.p + 3 * n;
        uint64_t capv = cap, outv = n;
        void* args[] = {(void*)&kp, (void*)&vp, (void*)&vp2, &capv, &ok, &ov, &ov2, &on, &outv};
        hip_check(hipModuleLaunchKernel(mod.cells_compact, blocks, 1, 1, 256, 1, 1, 0, s, args, nullptr),
                  "launch mbg_cells_compact");
        uint64_t got = 0;
        hip_check(hipMemcpyAsync(&got, on, 8, hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
        hip_check(hipMemcpyAsync(keys.data(), ok, n * 8, hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
        hip_check(hipMemcpyAsync(vals.data(), ov, n * 8, hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
        if (squares) hip_check(hipMemcpyAsync(vals2.data(), ov2, n * 8, hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
        hip_check(hipStreamSynchronize(s), "mbg_cells_compact");
        if (got != n) throw Error(MBG_ERR_HIP, "sparse window table: compaction count mismatch");
    }
    // counting sort on the window (W + 1 buckets), then each bucket by key
    std::vector<uint64_t> off((size_t)W + 2, 0);
    for (uint64_t i = 0; i < n; i++) off[(keys[i] >> K) + 1]++;
    for (int b = 0; b <= W; b++) off[b + 1] += off[b];
    // (key, value, square) sorted by key; without squares the third member is 0
    std::vector<std::tuple<uint64_t, uint64_t, uint64_t>> kv(n);
    {
        std::vector<uint64_t> pos(off.begin(), off.end() - 1);
        for (uint64_t i = 0; i < n; i++) kv[pos[keys[i] >> K]++] = {keys[i], vals[i], squares ? vals2[i] : 0};
    }
    const int nt = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    for (int t = 0; t < nt; t++)
        pool.emplace_back([&, t] {
            for (int b = t; b <= W; b += nt) std::sort(kv.begin() + (long)off[b], kv.begin() + (long)off[b + 1]);
        });
    for (auto& th : pool) th.join();
    const uint64_t smask = (1ull << K) - 1;
    const uint64_t nf = off[(size_t)W + 1] - off[(size_t)W], nc = n - nf;
    c->cell_w.resize(nc);
    c->cell_s.resize(nc);
    c->cell_v.resize(nc);
    c->cell_v2.assign(squares ? nc : 0, 0);
    c->fin_s.resize(nf);
    c->fin_v.resize(nf);
    for (uint64_t i = 0; i < nc; i++) {
        c->cell_w[i] = (uint32_t)(std::get<0>(kv[i]) >> K);
        c->cell_s[i] = std::get<0>(kv[i]) & smask;
        c->cell_v[i] = std::get<1>(kv[i]);
        if (squares) c->cell_v2[i] = std::get<2>(kv[i]);
    }
    for (uint64_t i = 0; i < nf; i++) {
        c->fin_s[i] = std::get<0>(kv[nc + i]) & smask;
        c->fin_v[i] = std::get<1>(kv[nc + i]);
    }
    c->cell_cap = n * 2 > cap ? std::min(cap * 2, kMaxCells) : cap;
}

// The shape a run uses without measurement: occupancy model plus tuning knobs.
// The lanes' state in HBM (MBG_STATE_HBM): occupancy is then set by registers (blocks of 256,
// about 4 waves per SIMD), and the tables stay in LDS when 4 blocks' copies fit beside the
// partial sums.
void make_state_hbm(Shape& sh, uint32_t blob_bytes, int block_req) {
    sh.state_hbm = 1;
    sh.block = block_req ? block_req : 256;
    sh.tables_lds = blob_bytes <= (36u << 10) ? 1 : 0;
}

Shape default_shape(const mbg_ctx* c, const mbg::CompiledModel& cm, const mbg::RateTables& rt, int block_req, int trace) {
    const bool rate32 = rates_fit_u32(rt);
    const bool compact = compact_fits(cm) && env_int("MBG_TUNE_COMPACT", 1) != 0;
    const uint32_t blob_bytes = pack_blob(cm, rt, rate32, compact).lay.bytes;
    // lanes' state in HBM by default where one wave's state would take more than half a CU's LDS
    // (one wave per CU in LDS; above about 106 words it would not fit at all)
    const int hbm_auto = lane_lds_bytes(cm.NW, 64, default_wsum_reg(cm.NW)) > (80u << 10) ? 1 : 0;
    int hbm_mode = env_int("MBG_TUNE_STATE_HBM", hbm_auto);
    const bool act_ok = act_hbm_allowed(default_wsum_reg(cm.NW), default_sumw(cm.NW, default_wsum_reg(cm.NW)));
    if (hbm_mode == 0 && !std::getenv("MBG_TUNE_STATE_HBM") && !block_req && act_ok) {
        // act alone in HBM by default where it keeps more waves resident (up to 12 per CU, which
        // leave the step loop its VGPRs): the 1,000-node network of 16 words runs 12 waves
        // instead of 8, 64 ms instead of 74 per million trajectories
        const Shape lds = choose_shape(c, cm, blob_bytes, 0, trace, rate32 ? 1 : 0, true);
        const size_t lds_bytes = lane_lds_bytes(cm.NW, lds.block, lds.wsum_reg) + (lds.tables_lds ? blob_bytes : 0);
        const int waves0 = std::min(16, (int)(160 * 1024 / lds_bytes) * (lds.block / 64));
        int waves2 = 0;
        const size_t th_bytes = (size_t)std::ceil(cm.max_time / cm.time_tick) * 8;
        act_hbm_block(c, cm.NW, (lds.tables_lds ? blob_bytes : 0) + th_bytes, 0, &waves2);
        if (waves2 > waves0) hbm_mode = 2;
    }
    const bool hbm = hbm_mode == 1 || hbm_mode == 3;
    // (act alone in HBM sets its own block below: the LDS-layout pass only picks the rest)
    const bool act_hbm = hbm_mode == 2 && act_ok;
    Shape sh = choose_shape(c, cm, blob_bytes, act_hbm ? 64 : block_req, trace, rate32 ? 1 : 0, !hbm);
    sh.compact = compact ? 1 : 0;
    sh.wsmad = sh.wsum_reg && rates_fit_i31(rt) && env_int("MBG_TUNE_WSMAD", 1) != 0 ? 1 : 0;
    sh.sparse = cm.K > mbg::kMaxReadouts ? 1 : 0;
    sh.prog = cm.any_prog ? 1 : 0;
    if (hbm) make_state_hbm(sh, blob_bytes, block_req);
    if (hbm_mode == 3 && sh.wsum_reg == 2) {
        // x alone in LDS, act and the word sums in HBM (asked for; the autotuner measures it)
        make_state_hbm(sh, blob_bytes, 0);
        sh.state_hbm = 3;
        sh.tables_lds = 0;
        const int breq = block_req ? block_req : env_int("MBG_TUNE_BLOCK", 256);
        const size_t th_bytes = (size_t)std::ceil(cm.max_time / cm.time_tick) * 8;
        int block = 0;
        for (int b = breq; b >= 64 && !block; b -= 64)
            if (lane_lds_bytes(cm.NW, b, 2, 3) + th_bytes <= (size_t)c->lds_per_block) block = b;
        sh.block = block ? block : 64;
    }
    if (act_hbm && act_hbm_allowed(sh.wsum_reg, sh.sumw)) {
        // act alone in HBM: x's LDS sets the block (a requested one if it fits), tables in LDS
        // where they fit beside it
        const int breq = block_req ? block_req : env_int("MBG_TUNE_BLOCK", 0);
        for (int tables = 1; tables >= 0; tables--) {
            int waves = 0;
            const size_t th_bytes = (size_t)std::ceil(cm.max_time / cm.time_tick) * 8;
            const int block = breq ? breq : act_hbm_block(c, cm.NW, (tables ? blob_bytes : 0) + th_bytes, 0, &waves);
            if (block && lane_lds_bytes(cm.NW, block, 1, 2) + (tables ? blob_bytes : 0) <= (size_t)c->lds_per_block) {
                sh.state_hbm = 2;
                sh.block = block;
                sh.tables_lds = tables;
                break;
            }
        }
        if (sh.state_hbm != 2 && block_req) {
            // as requested, but in the layout that fits it
            sh = choose_shape(c, cm, blob_bytes, block_req, trace, rate32 ? 1 : 0, true);
            sh.compact = compact ? 1 : 0;
            sh.wsmad = sh.wsum_reg && rates_fit_i31(rt) && env_int("MBG_TUNE_WSMAD", 1) != 0 ? 1 : 0;
            sh.sparse = cm.K > mbg::kMaxReadouts ? 1 : 0;
            sh.prog = cm.any_prog ? 1 : 0;
        }
    }
    return sh;
}

void run_impl(mbg_ctx* c, const mbg_model* m, const mbg_run_params* p, const Shape* force = nullptr) {
    const mbg::CompiledModel& cm = m->cm;
    if (cm.K > mbg::kMaxSparseReadouts) throw Error(MBG_ERR_LIMIT, mbg::kReadoutLimitMessage);
    if (!(cm.time_tick > 0.f) || !(cm.max_time > 0.f)) throw Error(MBG_ERR_ARG, "max_time and time_tick must be positive");
    const bool sparse = cm.K > mbg::kMaxReadouts;
    if (sparse && mbg::ceil_log2_u64((uint64_t)std::ceil(cm.max_time / cm.time_tick) + 1) + cm.K > 63)
        throw Error(MBG_ERR_LIMIT, "too many non-internal nodes for this number of windows (readouts + log2(windows + 1) "
                                   "must stay within 63 bits)");
    const uint64_t n = p->n_trajectories ? p->n_trajectories : (uint64_t)cm.sample_count;
    const uint64_t count = p->count ? p->count : n;
    if (n == 0 || count == 0) throw Error(MBG_ERR_ARG, "no trajectories to simulate");
    if (p->first_id + count > (1ull << 32)) throw Error(MBG_ERR_ARG, "trajectory ids must stay below 2^32");
    if (p->block && (p->block % 64 || p->block > 1024)) throw Error(MBG_ERR_ARG, "block must be a positive multiple of 64, at most 1024");

    const mbg::RateTables rt = mbg::make_rate_tables(cm, m->mult_up.empty() ? nullptr : m->mult_up.data(),
                                                     m->mult_down.empty() ? nullptr : m->mult_down.data(),
                                                     m->istate.empty() ? nullptr : m->istate.data());
    Shape sh;
    if (force) {
        sh = *force;
    } else {
        const auto tu = c->tuned.find(tune_key(cm, rt, p->trace));
        sh = (!p->block && tu != c->tuned.end()) ? tu->second : default_shape(c, cm, rt, p->block, p->trace);
    }
    sh.sparse = sparse ? 1 : 0;
    sh.prog = cm.any_prog ? 1 : 0;
    // MBG_TUNE_GENCODE=1: the fast nodes' logic as code generated for the model instead of
    // truth tables (a per-lane switch on the node id; measured several times slower, kept for
    // that comparison)
    sh.gen.clear();
    if (env_int("MBG_TUNE_GENCODE", 0) != 0 && m->parsed) sh.gen = generate_logic_source(*m->parsed, cm);
    const bool errors = p->error_columns != 0;
    sh.errors = errors ? 1 : 0;
    const Blob blob = pack_blob(cm, rt, sh.rate32 != 0, sh.compact != 0);
    const size_t state_smem = lane_lds_bytes(cm.NW, sh.block, sh.wsum_reg, sh.state_hbm) +
                              (sh.tables_lds ? blob.lay.bytes : 0);
    if (state_smem > (size_t)c->lds_per_block) throw Error(MBG_ERR_LIMIT, "launch shape does not fit in LDS");
    // Per-block partial sums in LDS, beside the shape's lanes and tables (they never change
    // the shape itself): the window entropies (W cells, summed by every trajectory in every
    // window: the hottest atomics of a run), and for small dense tables the whole window
    // table and the final states, as the reference does below 10 KB (window_average_small.cu:78).
    const size_t W_run = (size_t)std::ceil(cm.max_time / cm.time_tick);
    const size_t th_smem = W_run * 8 * (errors ? 2 : 1);  // with error columns, the squares too
    sh.th_lds = env_int("MBG_TUNE_TH_LDS", 1) != 0 && W_run <= 4096 &&
                        state_smem + th_smem <= (size_t)c->lds_per_block
                    ? 1
                    : 0;
    const size_t cells_smem = sparse ? 0 : (W_run + 1) * ((size_t)1 << cm.K) * 8;
    sh.cells_lds = sh.th_lds && !sparse && !errors && env_int("MBG_TUNE_CELLS_LDS", 1) != 0 && cells_smem <= (128u << 10) &&
                           state_smem + th_smem + cells_smem <= (size_t)c->lds_per_block
                       ? 1
                       : 0;
    const Module& mod = get_module(c, cm, sh);
    const size_t smem = state_smem + (sh.th_lds ? th_smem : 0) + (sh.cells_lds ? cells_smem : 0);
    int blocks_per_cu = 0;
    if (hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&blocks_per_cu, mod.sim, sh.block, smem) != hipSuccess ||
        blocks_per_cu < 1)
        blocks_per_cu = 1;

    const RunScales rs = run_scales(cm, n);
    const int W = rs.W;
    const uint64_t S = sparse ? 0 : rs.S;  // dense table only up to 20 readouts
    const int FP = rs.FP, FT = rs.FT;
    const uint64_t nsteps = rs.nsteps;
    // Fixed points: one slot per simulated trajectory, within a 16 GiB budget.
    const uint64_t fp_cap = std::min<uint64_t>(count, (16ull << 30) / (8ull * (uint64_t)cm.NW));
    // Initial states are staged per batch, within a 8 GiB budget.
    const uint64_t batch_max = std::max<uint64_t>(1, std::min<uint64_t>(count, (8ull << 30) / (8ull * stage_row(cm.NW))));

    hipStream_t s = c->stream;
    RunBuffers& B = c->buf;
    auto& d_blob = B.blob;
    auto& d_rtab = B.rtab;
    auto& d_stage = B.stage;
    auto& d_prob = B.prob;
    auto& d_th = B.th;
    auto& d_fin = B.fin;
    auto& d_fp = B.fp;
    auto& d_fpc = B.fpc;
    auto& d_summ = B.summ;
    auto& d_work = B.work;
    auto& d_flips = B.flips;
    auto& d_istate = B.istate;
    d_blob.upload(blob.words, s);
    d_rtab.upload(rt.rtab, s);
    d_istate.upload(rt.istate, s);
    d_stage.alloc((size_t)batch_max * stage_row(cm.NW));
    d_fp.alloc((size_t)fp_cap * cm.NW);
    const uint64_t summ_cap = p->trace ? count : 0;
    d_work.alloc(1);
    // Sparse table: one slot per distinct (window, projected state) cell and final state,
    // started at the size the last sparse run needed (at least 2^16, about 4 cells per
    // window per 256 trajectories) and grown four-fold whenever an insert finds no slot.
    uint64_t cell_cap = 0;
    if (sparse) {
        const uint64_t guess = 4ull * (uint64_t)(W + 1) * std::min<uint64_t>(count, 256);
        cell_cap = std::max<uint64_t>({1ull << 16, c->cell_cap, 1ull << mbg::ceil_log2_u64(guess)});
        cell_cap = std::min<uint64_t>(cell_cap, kMaxCells);
    }

    Desc D{};
    D.blob = d_blob.p;
    D.rtab = d_rtab.p;
    D.istate = d_istate.p;
    D.stage = d_stage.p;
    D.fp_states = d_fp.p;
    D.work_counter = d_work.p;
    // lanes' state in HBM: one column per thread of the largest grid a batch launches.  The
    // lanes' accesses are scattered 8-byte words, so the state is kept within about half of the
    // 256 MB Infinity Cache (fewer resident lanes, but cache hits instead of HBM lines): C's
    // 79-word lanes ran 307 ms at 65,536 lanes (124 MB) and 459 ms at 262,144 (497 MB).
    uint64_t grid_max = p->grid ? (uint64_t)p->grid : (uint64_t)c->cus * (uint64_t)blocks_per_cu;
    if (sh.state_hbm && !p->grid) {
        const uint64_t lane_bytes = (uint64_t)lane_hbm_words(cm.NW, sh.wsum_reg, sh.state_hbm) * 8;
        const uint64_t fit = (uint64_t)env_int("MBG_TUNE_STATE_MB", 128) * (1ull << 20) / (lane_bytes * (uint64_t)sh.block);
        grid_max = std::max<uint64_t>((uint64_t)c->cus, std::min<uint64_t>(grid_max, fit));
    }
    if (sh.state_hbm) {
        const uint64_t stride = grid_max * (uint64_t)sh.block;
        if (stride > 0xffffffffull / (3ull * (uint64_t)cm.NW)) throw Error(MBG_ERR_LIMIT, "grid too large for the lanes' state");
        B.lane_state.alloc((size_t)stride * lane_hbm_words(cm.NW, sh.wsum_reg, sh.state_hbm));
        D.lane_state = B.lane_state.p;
        D.lane_stride = (uint32_t)stride;
    }
    D.fp_cap = fp_cap;
    D.summ_cap = summ_cap;
    D.nsteps = nsteps;
    D.first_id = p->first_id;
    D.seed = p->use_model_seed ? (uint64_t)cm.seed : p->seed;
    D.maxt = (double)cm.max_time;
    D.tick = (double)cm.time_tick;
    D.inv_tick = 1.0 / D.tick;
    D.rscale = std::ldexp(1.0, -rt.E);
    D.lay = blob.lay;
    D.variant_id = p->variant_id;
    D.N = cm.N;
    D.K = cm.K;
    D.W = W;
    D.FP = FP;
    D.FT = FT;
    D.sigma = rt.sigma;
    D.compat = p->compat_batch_resume ? 1 : 0;
    // Error columns: each lane lists its current window's (state, time) entries in a lane-
    // interleaved buffer (one column per thread of the largest grid), err_cap entries long to
    // start with (what the last run needed) and four times longer after an overflow; the run is
    // deterministic, so the repeat gives the same tables.
    uint32_t err_cap = errors ? c->err_cap : 0;
    const uint64_t err_stride = grid_max * (uint64_t)sh.block;
    if (errors) {
        if (err_stride > 0xffffffffull) throw Error(MBG_ERR_LIMIT, "grid too large for the error columns");
        const RunScales rsq = run_scales(cm, n);
        D.FQ = rsq.FQ;
        D.FQT = rsq.FQT;
        D.sq_scale = rsq.sq_scale;
        D.sq_scale_th = rsq.sq_scale_th;
        D.err_stride = (uint32_t)err_stride;
    }

    EventPair ev;
    hipEvent_t e0 = ev.a, e1 = ev.b;
    float sim_ms = 0.f;
    for (;;) {
        d_prob.zero((size_t)W * S, s);
        d_th.zero((size_t)W, s);
        d_fin.zero((size_t)S, s);
        d_fpc.zero(1, s);
        d_summ.zero((size_t)summ_cap * 4, s);
        d_flips.zero(1, s);
        if (sparse) {
            B.cells.fill(cell_cap, 0xff, s);  // every key ~0: empty
            B.cell_vals.zero(cell_cap, s);
            B.cell_miss.zero(1, s);
        }
        if (errors) {
            B.prob2.zero((size_t)W * S, s);
            B.th2.zero((size_t)W, s);
            if (sparse) B.cell_vals2.zero(cell_cap, s);
            B.err_s.alloc((size_t)err_cap * err_stride);
            B.err_v.alloc((size_t)err_cap * err_stride);
            B.err_miss.zero(1, s);
            D.prob2 = B.prob2.p;
            D.th2 = B.th2.p;
            D.cell_vals2 = sparse ? B.cell_vals2.p : nullptr;
            D.err_s = B.err_s.p;
            D.err_v = B.err_v.p;
            D.err_miss = B.err_miss.p;
            D.err_cap = err_cap;
        }
        D.prob = d_prob.p;
        D.th = d_th.p;
        D.fin = d_fin.p;
        D.fp_count = d_fpc.p;
        D.summ = d_summ.p;
        D.flip_total = d_flips.p;
        D.cells = sparse ? B.cells.p : nullptr;
        D.cell_vals = sparse ? B.cell_vals.p : nullptr;
        D.cell_miss = sparse ? B.cell_miss.p : nullptr;
        D.cell_mask = sparse ? cell_cap - 1 : 0;
        check_desc(D, sparse, sh.state_hbm != 0);
        sim_ms = 0.f;
        for (uint64_t done = 0; done < count; done += D.batch) {
            D.batch = std::min(batch_max, count - done);
            D.batch_first = p->first_id + done;
            hip_check(hipMemsetAsync(d_work.p, 0, 8, s), "hipMemsetAsync");
            const uint64_t init_threads = D.batch * (uint64_t)cm.NW;
            // mbg_init_rates: one block per 64 trajectories (per (trajectory, word) above 128 words)
            const uint64_t rate_blocks = cm.NW <= 128 ? (D.batch + 63) / 64 : (init_threads + 255) / 256;
            if (cm.NW > 128) launch(mod.init_state, (init_threads + 255) / 256, 256, 0, s, D, "launch mbg_init_state");
            launch(mod.init_rates, rate_blocks, 256, 0, s, D, "launch mbg_init_rates");
            if (cm.NW > 128) launch(mod.init_readouts, (D.batch + 255) / 256, 256, 0, s, D, "launch mbg_init_readouts");
            uint64_t grid = grid_max;
            grid = std::max<uint64_t>(1, std::min<uint64_t>(grid, (D.batch + sh.block - 1) / sh.block));
            if (sh.state_hbm && grid * (uint64_t)sh.block > (uint64_t)D.lane_stride)
                throw Error(MBG_ERR_ARG, "internal error: the grid's lanes do not fit the lane-state buffer");
            hip_check(hipEventRecord(e0, s), "hipEventRecord");
            launch(mod.sim, grid, sh.block, smem, s, D, "launch mbg_simulate");
            hip_check(hipEventRecord(e1, s), "hipEventRecord");
            hip_check(hipStreamSynchronize(s), "kernel execution");

            float ms = 0.f;
            hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
            sim_ms += ms;
        }
        bool again = false;
        if (sparse) {
            uint64_t miss = 0;
            hip_check(hipMemcpy(&miss, B.cell_miss.p, 8, hipMemcpyDeviceToHost), "hipMemcpy");
            if (miss) {
                if (cell_cap >= kMaxCells)
                    throw Error(MBG_ERR_LIMIT, "sparse window table full: more distinct (window, state) cells than " +
                                                   std::to_string(kMaxCells / 2));
                cell_cap = std::min<uint64_t>(cell_cap * 4, kMaxCells);  // repeat the run with a larger table
                again = true;
            }
        }
        if (errors) {
            uint64_t miss = 0;
            hip_check(hipMemcpy(&miss, B.err_miss.p, 8, hipMemcpyDeviceToHost), "hipMemcpy");
            if (miss) {
                if (err_cap >= 4096)
                    throw Error(MBG_ERR_LIMIT, "error columns: a trajectory visits more than 4,096 projected states in one window");
                err_cap *= 4;  // repeat the run with longer window lists
                again = true;
            }
        }
        if (!again) break;
    }
    if (errors) c->err_cap = err_cap;

    c->prob.resize((size_t)W * S);  // filled from the device by the first getter that needs it
    c->prob_on_host = false;
    c->th.resize(W);
    c->fin.resize(S);
    hip_check(hipMemcpy(c->th.data(), d_th.p, c->th.size() * 8, hipMemcpyDeviceToHost), "hipMemcpy");
    if (S) hip_check(hipMemcpy(c->fin.data(), d_fin.p, c->fin.size() * 8, hipMemcpyDeviceToHost), "hipMemcpy");
    c->sparse = sparse;
    if (sparse) read_cells(c, mod, cell_cap, W, cm.K, s, errors);
    c->prob2.clear();
    c->th2.clear();
    if (!sparse) c->cell_v2.clear();
    if (errors) {
        c->th2.resize(W);
        hip_check(hipMemcpy(c->th2.data(), B.th2.p, (size_t)W * 8, hipMemcpyDeviceToHost), "hipMemcpy");
        if (!sparse) {
            c->prob2.resize((size_t)W * S);
            hip_check(hipMemcpy(c->prob2.data(), B.prob2.p, c->prob2.size() * 8, hipMemcpyDeviceToHost), "hipMemcpy");
        }
    }
    uint64_t nfp = 0, flips = 0;
    hip_check(hipMemcpy(&nfp, d_fpc.p, 8, hipMemcpyDeviceToHost), "hipMemcpy");
    hip_check(hipMemcpy(&flips, d_flips.p, 8, hipMemcpyDeviceToHost), "hipMemcpy");
    if (nfp > fp_cap) throw Error(MBG_ERR_LIMIT, "fixed-point buffer overflow");
    c->summ.resize((size_t)summ_cap * 4);
    if (summ_cap) hip_check(hipMemcpy(c->summ.data(), d_summ.p, c->summ.size() * 8, hipMemcpyDeviceToHost), "hipMemcpy");
    // distinct fixed points in the reference order (src/state.h:43-49)
    collect_fixed_points(c, mod, d_fp.p, nfp, cm.NW, cm.NW, s, c->fixed_states, c->fixed_counts);

    mbg_run_info& I = c->info;
    I.windows = W;
    I.readouts = cm.K;
    I.prob_exponent = FP;
    I.th_exponent = FT;
    I.rate_exponent = rt.E;
    I.state_words = cm.NW;
    I.discrete = cm.discrete ? 1 : 0;
    I.n_trajectories = n;
    I.simulated = count;
    I.transitions = flips;
    I.fixed_point_trajectories = nfp;
    I.kernel_ms = sim_ms;
    I.time_tick = cm.time_tick;
    I.max_time = cm.max_time;
    I.block = sh.block;
    I.tables_in_lds = sh.tables_lds;
    I.lds_bytes = (int)smem;
    I.blocks_per_cu = blocks_per_cu;
    I.state_hbm = sh.state_hbm;
    I.error_columns = errors ? 1 : 0;
    I.sq_exponent = errors ? D.FQ : 0;
    I.th_sq_exponent = errors ? D.FQT : 0;
    c->marg_fn = mod.marginals;
    c->have = true;
}

// Measured launch shape for a model: coordinate descent over the layout (table placement,
// compact tables, register or LDS rate sums, a lower-occupancy block) and then the step
// knobs (dependents per round, search width, multiply-add sum updates), each candidate timed
// on `sample` trajectories (HIP events around the kernel).  Results do not depend on the shape
// (every draw is keyed by trajectory id and every sum is exact), so only speed is compared.
// The fastest shape is kept for later runs of models with the same structure.
void autotune_impl(mbg_ctx* c, const mbg_model* m, uint64_t sample, int trace, mbg_tune_info* out) {
    const mbg::CompiledModel& cm = m->cm;
    const mbg::RateTables rt = mbg::make_rate_tables(cm, m->mult_up.empty() ? nullptr : m->mult_up.data(),
                                                     m->mult_down.empty() ? nullptr : m->mult_down.data(),
                                                     m->istate.empty() ? nullptr : m->istate.data());
    const bool r32 = rates_fit_u32(rt), r31 = rates_fit_i31(rt), cfit = compact_fits(cm);
    // default sample: about four trajectories per lane of a full chip (1,024 lanes per CU at
    // most), so that the run's tail (lanes out of work while others finish) weighs as it does in
    // a run of millions; with one or two per lane the tail decided the dependents cap
    const uint64_t n = sample ? sample
                              : std::min<uint64_t>((uint64_t)cm.sample_count,
                                                   std::max<uint64_t>(262144, 4ull * (uint64_t)c->cus * 1024ull));
    mbg_run_params rp{};
    rp.n_trajectories = n;
    rp.count = n;
    rp.use_model_seed = 1;
    rp.trace = trace;
    int trials = 0;
    std::set<Shape> tried;
    auto measure = [&](const Shape& sh) -> double {
        if (!tried.insert(sh).second) return 1e300;
        trials++;
        try {
            run_impl(c, m, &rp, &sh);
        } catch (const Error& e) {
            if (e.code != MBG_ERR_LIMIT) throw;
            return 1e300;  // does not fit
        }
        return c->info.kernel_ms;
    };
    const Shape base = default_shape(c, cm, rt, 0, trace);
    Shape best = base;
    const double base_ms = measure(base);
    double best_ms = base_ms;
    // a candidate replaces the best only when it is more than 1 % faster: smaller differences
    // are within the sample's run-to-run spread, and the untuned default is the better-studied
    // shape on the full run
    auto consider = [&](const Shape& sh) {
        const double ms = measure(sh);
        if (ms < best_ms * 0.99) {
            best_ms = ms;
            best = sh;
        }
    };
    // layout: each knob of the default toggled, then one occupancy step down
    {
        std::vector<Shape> cands;
        auto add = [&](int compact, int wreg, int tables, int below) {
            if (wreg == 1 && !reg_sums_fit(cm.NW)) return;
            if (compact && !cfit) return;
            const uint32_t blob = pack_blob(cm, rt, r32, compact != 0).lay.bytes;
            int waves = 0;
            const int block = layout_block(c, cm.NW, wreg, blob, tables, &waves, below);
            if (!block) return;
            Shape sh = base;
            sh.state_hbm = 0;  // these layouts keep the lanes' whole state in LDS
            sh.compact = compact;
            sh.wsum_reg = wreg;
            sh.sumw = default_sumw(cm.NW, wreg);
            sh.wsmad = wreg && r31 ? 1 : 0;
            sh.tables_lds = tables;
            sh.block = block;
            cands.push_back(sh);
        };
        if (base.state_hbm == 2) add(base.compact, base.wsum_reg, base.tables_lds, 1 << 30);
        add(base.compact, base.wsum_reg, !base.tables_lds, 1 << 30);
        add(!base.compact, base.wsum_reg, base.tables_lds, 1 << 30);
        add(base.compact, base.wsum_reg ? 0 : (reg_sums_fit(cm.NW) ? 1 : 2), base.tables_lds, 1 << 30);
        // 17-32 words: the two-level sums search one word instead of a two-word group
        if (base.wsum_reg == 1 && cm.NW > 16) add(base.compact, 2, base.tables_lds, 1 << 30);
        {
            const uint32_t blob = pack_blob(cm, rt, r32, base.compact != 0).lay.bytes;
            const size_t lds = lane_lds_bytes(cm.NW, base.block, base.wsum_reg) + (base.tables_lds ? blob : 0);
            const int waves = std::min(16, (int)(160 * 1024 / lds) * (base.block / 64));
            add(base.compact, base.wsum_reg, base.tables_lds, waves);
        }
        for (const Shape& sh : cands) consider(sh);
        // the lanes' state in HBM, tables as the default places them
        Shape h = base;
        make_state_hbm(h, pack_blob(cm, rt, r32, base.compact != 0).lay.bytes, 0);
        consider(h);
        // x alone in LDS, act and the word sums in HBM (two-level sums, tables in HBM): the
        // dependents' input gathers and the flips read LDS instead of the HBM buffer; the
        // largest block whose x fits beside the window partial sums, and blocks of 256
        if (base.wsum_reg == 2) {
            const size_t th_bytes = (size_t)std::ceil(cm.max_time / cm.time_tick) * 8;
            int prev = 0;
            for (int cap : {1024, 256}) {
                int block = 0;
                for (int b = cap; b >= 64 && !block; b -= 64)
                    if (lane_lds_bytes(cm.NW, b, 2, 3) + th_bytes <= (size_t)c->lds_per_block) block = b;
                if (!block || block == prev) continue;
                prev = block;
                Shape x = h;
                x.state_hbm = 3;
                x.tables_lds = 0;
                x.block = block;
                consider(x);
            }
        }
        // act alone in HBM (x in LDS, word sums in VGPRs): blocks of at most 12 waves (3 per
        // SIMD leave the step loop its ~170 VGPRs without spills), and the fullest block
        if (act_hbm_allowed(base.wsum_reg, base.sumw)) {
            const uint32_t blob = pack_blob(cm, rt, r32, base.compact != 0).lay.bytes;
            const size_t th_bytes = (size_t)std::ceil(cm.max_time / cm.time_tick) * 8;
            int prev = 0;
            for (int full = 0; full < 2; full++) {
                int waves = 0;
                const int block = full ? layout_block(c, cm.NW, 1, blob, base.tables_lds, &waves, 1 << 30, 2)
                                       : act_hbm_block(c, cm.NW, (base.tables_lds ? blob : 0) + th_bytes, 0, &waves);
                if (!block || block == prev) continue;
                prev = block;
                Shape a = base;
                a.state_hbm = 2;
                a.block = block;
                consider(a);
            }
        }
    }
    // step knobs on the best layout
    {
        const Shape lay = best;
        // dependents per iteration: no cap, or a few rounds (one dependent per round, see
        // dependents_per_round)
        {
            const Shape b = best;
            for (int rounds : {0, 2, 3, 4, 6}) {
                Shape sh = b;
                sh.dep_cap = rounds ? std::max(2, rounds * b.dep_batch) : 0;
                consider(sh);
            }
        }
        Shape sh = lay;
        sh.scan = lay.scan == 8 ? 4 : 8;
        consider(sh);
        if (best.wsum_reg && r31) {
            Shape w = best;
            w.wsmad = !best.wsmad;
            consider(w);
        }
    }
    c->tuned[tune_key(cm, rt, trace)] = best;
    c->have = false;
    if (out) {
        *out = mbg_tune_info{};
        out->block = best.block;
        out->tables_in_lds = best.tables_lds;
        out->wsum_reg = best.wsum_reg;
        out->compact = best.compact;
        out->dep_batch = best.dep_batch;
        out->scan = best.scan;
        out->wsmad = best.wsmad;
        out->trials = trials;
        out->sample = n;
        out->best_ms = best_ms;
        out->default_ms = base_ms;
        out->dep_cap = best.dep_cap;
        out->state_hbm = best.state_hbm;
    }
}

// State-word class of a sweep variant: 1, 2, 4, 8, then multiples of 4 words up to 32, then
// powers of two.  Lanes of a class hold NWc words, so a tight class keeps more lanes in LDS.
int word_class(int nw) {
    if (nw <= 8) {
        int c = 1;
        while (c < nw) c *= 2;
        return c;
    }
    if (nw <= 32) return (nw + 3) / 4 * 4;
    int c = 64;
    while (c < nw) c *= 2;
    return c;
}

std::vector<uint64_t> variant_key(const mbg_variant& v) {
    return {(uint64_t)(uintptr_t)v.model, v.n_trajectories, v.use_model_seed ? ~0ull : v.seed, v.variant_id,
            v.model ? (uint64_t)v.model->cm.N : 0ull};
}

// Variant sweep: all variants in one HIP graph launch.
void run_variants_impl(mbg_ctx* c, const mbg_variant* vs, int nv, int flags, int depth = 0) {
    const bool errors = (flags & MBG_SWEEP_ERROR_COLUMNS) != 0;
    const auto t0 = std::chrono::steady_clock::now();
    if (c->sweep) {
        // return the previous sweep's buffers to the pool, last variant first, so that a sweep of
        // the same variants takes each one's buffers back at the sizes it needs (no reallocation)
        for (auto it = c->sweep->vars.rbegin(); it != c->sweep->vars.rend(); ++it)
            c->variant_pool.push_back(std::move(*it));
        c->sweep.reset();
    }
    auto sw = std::make_unique<SweepState>();
    hipStream_t main = c->stream;
    std::map<std::pair<int, int>, int> class_of;  // (words, discrete) -> class
    std::vector<mbg::RateTables> rts(nv);
    std::vector<int> fits32(nv);
    for (int v = 0; v < nv; v++) {
        const mbg_model* m = vs[v].model;
        if (!m) throw Error(MBG_ERR_ARG, "variant without a model");
        const mbg::CompiledModel& cm = m->cm;
        if (cm.K > mbg::kMaxReadouts)
            throw Error(MBG_ERR_LIMIT, "variant sweeps support at most 20 non-internal nodes (run the variant alone)");
        if (!(cm.time_tick > 0.f) || !(cm.max_time > 0.f)) throw Error(MBG_ERR_ARG, "max_time and time_tick must be positive");
        std::unique_ptr<VariantRun> V;
        if (!c->variant_pool.empty()) {
            V = std::move(c->variant_pool.back());
            c->variant_pool.pop_back();
            V->fixed_states.clear();
            V->fixed_counts.clear();
        } else {
            V = std::make_unique<VariantRun>();
        }
        V->m = m;
        V->n = vs[v].n_trajectories ? vs[v].n_trajectories : (uint64_t)cm.sample_count;
        if (V->n == 0 || V->n > (1ull << 32)) throw Error(MBG_ERR_ARG, "invalid trajectory count");
        const int nwc = word_class(cm.NW);
        V->NWc = nwc;
        rts[v] = mbg::make_rate_tables(cm, m->mult_up.empty() ? nullptr : m->mult_up.data(),
                                       m->mult_down.empty() ? nullptr : m->mult_down.data(),
                                       m->istate.empty() ? nullptr : m->istate.data());
        fits32[v] = rates_fit_u32(rts[v]);
        const auto key = std::make_pair(nwc, cm.discrete ? 1 : 0);
        auto it = class_of.find(key);
        if (it == class_of.end()) {
            it = class_of.emplace(key, (int)sw->classes.size()).first;
            sw->classes.push_back(std::make_unique<SweepClass>());
            Shape& sh = sw->classes.back()->sh;
            sh = Shape{nwc, 64, cm.discrete ? 1 : 0, 0, 0, 0, 1};
            sh.rate32 = 1;
            sh.multi = 1;
            sh.compact = env_int("MBG_TUNE_COMPACT", 1) != 0 ? 1 : 0;
            sh.wsmad = env_int("MBG_TUNE_WSMAD", 1) != 0 ? 1 : 0;
            sh.wsum_reg = default_wsum_reg(nwc);
            sh.sumw = default_sumw(nwc, sh.wsum_reg);
            sh.errors = errors ? 1 : 0;
            sh.sweep_xcd = std::max(0, std::min(2, env_int("MBG_TUNE_SWEEP_XCD", 2)));
        }
        SweepClass& C = *sw->classes[it->second];
        C.members.push_back(v);
        C.work += (double)V->n * (double)cm.N * (double)cm.max_time;
        C.sh.generic |= cm.any_generic ? 1 : 0;
        C.sh.prog |= cm.any_prog ? 1 : 0;
        C.sh.max_in = std::max(C.sh.max_in, cm.max_inputs);
        C.sh.rate32 &= fits32[v];
        C.sh.compact &= compact_fits(cm) ? 1 : 0;
        C.sh.wsmad &= rates_fit_i31(rts[v]) ? 1 : 0;
        sw->vars.push_back(std::move(V));
    }
    // per class: the variants in decreasing order of their expected work per trajectory, the
    // order in which the class kernel's waves drain them (the run ends on the cheapest
    // trajectories, so its tail is short) and the share of blocks each gets with the tables in
    // LDS.  The expectation is the transitions per trajectory an earlier sweep of the variant
    // measured, else a rate estimate: the mean of every node's two rates times max_time (which
    // misses the active fraction: 30-45 % off from variant to variant on random networks)
    std::vector<double> member_cost(nv, 0.0);
    std::list<std::vector<int>> keep_starts;  // host sources of the block-start uploads
    for (auto& Cp : sw->classes) {
        std::vector<double>& key = member_cost;
        double ratio = 0.0;  // measured / estimated, summed over the members measured before
        int measured = 0;
        std::vector<char> known(nv, 0);
        for (int v : Cp->members) {
            double sum = 0.0;
            for (uint64_t r : rts[v].rates) sum += (double)r;
            const double est = std::ldexp(sum, -rts[v].E) * (double)sw->vars[v]->m->cm.max_time;
            const auto mc = c->variant_cost.find(variant_key(vs[v]));
            if (mc != c->variant_cost.end() && est > 0) {
                key[v] = mc->second;
                ratio += mc->second / est;
                measured++;
                known[v] = 1;
            } else {
                key[v] = est;
            }
        }
        // estimates on the scale of the measured members (when some are)
        if (measured)
            for (int v : Cp->members)
                if (!known[v]) key[v] *= ratio / measured;
        std::stable_sort(Cp->members.begin(), Cp->members.end(), [&](int a, int b) { return key[a] > key[b]; });
    }
    // per class: block size for the lanes' LDS, module, occupancy
    const size_t cu_lds = 160 * 1024;
    for (auto& Cp : sw->classes) {
        SweepClass& C = *Cp;
        // tables in LDS: every wave of a block works on the block's current variant, whose tables
        // the block stages beside the lanes' state (the class's largest blob, plus a 16-byte flag)
        C.tab_bytes = 0;
        for (int v : C.members)
            C.tab_bytes = std::max(C.tab_bytes, pack_blob(sw->vars[v]->m->cm, rts[v], C.sh.rate32 != 0,
                                                          C.sh.compact != 0).lay.bytes);
        // the class's key: its shape so far and the structure of its members
        C.key = {C.sh.nw, C.sh.discrete, C.sh.errors, C.sh.generic, C.sh.prog, C.sh.max_in, (long long)C.members.size()};
        for (int v : C.members) C.key.push_back(sw->vars[v]->m->cm.N);
        ClassChoice ch;
        if (c->force_class) {
            ch = *c->force_class;
        } else {
            const auto tc = c->tuned_classes.find(C.key);
            if (tc != c->tuned_classes.end()) {
                ch = tc->second;
                C.tuned = true;
            }
        }
        // default: the tables in LDS up to 56 KiB (measured on the 256-variant sweep: every class
        // up to 28 words faster, by 10-30 %; the 29-32-word class, 56-64 KiB of tables, as fast
        // with them in HBM at more resident waves)
        const int want_tables = ch.tables_lds >= 0 ? ch.tables_lds
                                                   : env_int("MBG_TUNE_SWEEP_TABLES", C.tab_bytes <= (56u << 10) ? 1 : 0);
        C.sh.tables_lds = want_tables && C.tab_bytes <= (64u << 10) ? 1 : 0;
        const size_t fixed = C.sh.tables_lds ? (size_t)C.tab_bytes + 16 : 0;
        int best_waves = 0;
        for (int block = 1024; block >= 64; block -= 64) {
            const size_t lds = lane_lds_bytes(C.sh.nw, block, C.sh.wsum_reg) + fixed;
            if (lds > (size_t)c->lds_per_block) continue;
            const int waves = std::min(16, (int)(cu_lds / lds) * (block / 64));
            if (waves > best_waves) {
                best_waves = waves;
                C.sh.block = block;
            }
        }
        if (!best_waves) throw Error(MBG_ERR_LIMIT, "model too large: per-lane state does not fit in LDS");
        if (const int bq = env_int("MBG_TUNE_SWEEP_BLOCK", 0)) {
            if (bq % 64 == 0 && bq <= 1024 && lane_lds_bytes(C.sh.nw, bq, C.sh.wsum_reg) + fixed <= (size_t)c->lds_per_block)
                C.sh.block = bq;
        }
        // the active sets in HBM where that keeps more waves resident (up to 12 per CU, which
        // leave the step loop its VGPRs), as in single runs
        // leave room for the per-wave window-entropy partials
        if (act_hbm_allowed(C.sh.wsum_reg, C.sh.sumw) &&
            (ch.state_hbm >= 0 ? ch.state_hbm == 2 : env_int("MBG_TUNE_STATE_HBM", -1) != 0)) {
            int wm = 1;
            for (int v : C.members) {
                const mbg::CompiledModel& cmv = sw->vars[v]->m->cm;
                wm = std::max(wm, (int)std::ceil(cmv.max_time / cmv.time_tick));
            }
            int waves2 = 0;
            const int block2 = act_hbm_block(c, C.sh.nw, fixed, (size_t)wm * 8 * (errors ? 2 : 1), &waves2);
            if (block2 && (waves2 > best_waves || ch.state_hbm == 2)) {
                C.sh.state_hbm = 2;
                C.sh.block = block2;
            }
        }
        // the whole lane state in HBM (one buffer per class)
        if (env_int("MBG_TUNE_SWEEP_STATE_HBM", 0) == 1 && C.sh.state_hbm != 2) {
            C.sh.state_hbm = 1;
            C.sh.block = 256;
        }
        // per-wave window-entropy partials in LDS when they fit without costing occupancy
        int wmax = 1;
        for (int v : C.members) {
            const mbg::CompiledModel& cmv = sw->vars[v]->m->cm;
            wmax = std::max(wmax, (int)std::ceil(cmv.max_time / cmv.time_tick));
        }
        const size_t state_lds = lane_lds_bytes(C.sh.nw, C.sh.block, C.sh.wsum_reg, C.sh.state_hbm) + fixed;
        const size_t th_lds = (size_t)(C.sh.block / 64) * (size_t)wmax * 8 * (errors ? 2 : 1);  // + squares
        // (with the state in HBM: when 16 waves' slices fit in a CU's LDS)
        const bool th_fits = state_lds ? cu_lds / (state_lds + th_lds) == cu_lds / state_lds
                                       : th_lds * (size_t)(16 / (C.sh.block / 64)) <= cu_lds;
        C.sh.th_lds = env_int("MBG_TUNE_TH_LDS", 1) != 0 && wmax <= 4096 &&
                              state_lds + th_lds <= (size_t)c->lds_per_block && th_fits
                          ? 1
                          : 0;
        C.wmax = wmax;
        C.sh.dep_batch = dependents_per_round();
        C.sh.dep_cap = 0;
        for (int v : C.members)
            C.sh.dep_cap = std::max(C.sh.dep_cap, default_dep_cap(sw->vars[v]->m->cm, C.sh.dep_batch));
        if (ch.dep_cap >= 0) C.sh.dep_cap = ch.dep_cap;
        C.sh.scan = env_int("MBG_TUNE_SCAN", 4) >= 8 ? 8 : 4;
        const Module& mod = get_module(c, sw->vars[C.members[0]]->m->cm, C.sh);
        for (int v : C.members) sw->vars[v]->mod = &mod;
        C.sim = mod.sim;
        C.init_state = mod.init_state;
        C.init_rates = mod.init_rates;
        C.init_readouts = mod.init_readouts;
        C.marginals = mod.marginals;
        C.marginals_batch = mod.marginals_batch;
        C.init_rates_multi = mod.init_rates_multi;
        C.smem = state_lds + (C.sh.th_lds ? th_lds : 0);
        int nb = 0;
        if (hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&nb, C.sim, C.sh.block, C.smem) != hipSuccess || nb < 1) nb = 1;
        if (C.sh.state_hbm == 1) {
            // keep the class's lane state within about half of the Infinity Cache
            const uint64_t lane_bytes = (uint64_t)(C.sh.wsum_reg == 1 ? 2 : 3) * (uint64_t)C.sh.nw * 8;
            const uint64_t fit = (uint64_t)env_int("MBG_TUNE_STATE_MB", 128) * (1ull << 20) /
                                 (lane_bytes * (uint64_t)C.sh.block * (uint64_t)c->cus);
            nb = (int)std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)nb, fit));
        }
        C.blocks_per_cu = nb;
        {
            uint64_t traj = 0;
            for (int v : C.members) traj += sw->vars[v]->n;
            C.grid = std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)c->cus * (uint64_t)nb,
                                                              (traj + C.sh.block - 1) / C.sh.block));
        }
        if (C.sh.tables_lds) {
            // block b starts at the member that holds the point (b + 1/2) / grid of the class's
            // work (trajectories x expected work per trajectory, members in decreasing order of
            // the latter), so the blocks on a variant are in proportion to its work
            std::vector<double> cum;
            double total = 0.0;
            for (int v : C.members) {
                total += (double)sw->vars[v]->n * std::max(member_cost[v], 1e-30);
                cum.push_back(total);
            }
            std::vector<int>& bs = keep_starts.emplace_back(C.grid, 0);
            for (uint64_t b = 0; b < C.grid; b++) {
                const double pos = ((double)b + 0.5) / (double)C.grid * total;
                const size_t k = (size_t)(std::upper_bound(cum.begin(), cum.end(), pos) - cum.begin());
                bs[b] = (int)std::min(k, cum.size() - 1);
            }
            C.block_start.upload(bs, main);
        }
        if (C.sh.state_hbm)
            C.act.alloc((size_t)c->cus * (size_t)nb * (size_t)C.sh.block * (size_t)C.sh.nw *
                        (C.sh.state_hbm == 2 ? 1 : C.sh.wsum_reg == 1 ? 2 : 3));
        if (errors) {
            const size_t cols = (size_t)c->cus * (size_t)nb * (size_t)C.sh.block;
            C.err_s.alloc((size_t)c->err_cap * cols);
            C.err_v.alloc((size_t)c->err_cap * cols);
            C.err_miss.zero(1, main);
        }
    }
    // shared per-sweep buffers: counters and the marginal jobs of every variant
    SweepShared& SS = c->sweep_shared;
    SS.counters.zero((size_t)2 * nv, main);
    std::vector<MargJob> jobs;
    std::vector<int> job_of(nv, -1);
    uint64_t marg_total = 0;
    int w_max = 0;
    // per variant: tables and buffers.  The host sources of the asynchronous uploads (packed
    // tables, descriptors) are kept until the uploads have completed: a copy from pageable
    // memory may read its source after hipMemcpyAsync returns.
    std::vector<std::vector<uint64_t>> keep_blobs;
    std::vector<std::vector<Desc>> keep_descs;
    keep_blobs.reserve((size_t)nv);
    keep_descs.reserve(sw->classes.size());
    for (auto& Cp : sw->classes) {
        SweepClass& C = *Cp;
        keep_descs.emplace_back();
        std::vector<Desc>& descs = keep_descs.back();
        for (int v : C.members) {
            VariantRun& V = *sw->vars[v];
            const mbg::CompiledModel& cm = V.m->cm;
            const mbg::RateTables& rt = rts[v];
            Blob blob = pack_blob(cm, rt, C.sh.rate32 != 0, C.sh.compact != 0);
            const RunScales rs = run_scales(cm, V.n);
            const uint64_t fp_cap = V.n;
            RunBuffers& B = V.buf;
            B.blob.upload(blob.words, main);
            keep_blobs.push_back(std::move(blob.words));
            B.rtab.upload(rt.rtab, main);
            B.istate.upload(rt.istate, main);
            B.stage.alloc((size_t)V.n * stage_row(V.NWc));
            B.prob.zero((size_t)rs.W * rs.S, main);
            B.th.zero((size_t)rs.W, main);
            B.fin.zero((size_t)rs.S, main);
            B.fp.alloc((size_t)fp_cap * V.NWc);
            B.work.zero(1, main);
            if (errors) {
                B.prob2.zero((size_t)rs.W * rs.S, main);
                B.th2.zero((size_t)rs.W, main);
            }
            if (cm.K > 0) {
                job_of[v] = (int)jobs.size();
                jobs.push_back(MargJob{B.prob.p, marg_total, rs.W, cm.K});
                marg_total += (uint64_t)rs.W * (uint64_t)cm.K;
                w_max = std::max(w_max, rs.W);
            }
            Desc& D = V.D;
            D = Desc{};
            D.blob = B.blob.p;
            D.rtab = B.rtab.p;
            D.istate = B.istate.p;
            D.stage = B.stage.p;
            D.prob = B.prob.p;
            D.th = B.th.p;
            D.fin = B.fin.p;
            D.fp_states = B.fp.p;
            D.fp_count = SS.counters.p + 2 * (size_t)v;
            D.summ = nullptr;
            D.work_counter = B.work.p;
            D.flip_total = SS.counters.p + 2 * (size_t)v + 1;
            D.fp_cap = fp_cap;
            D.summ_cap = 0;
            D.nsteps = rs.nsteps;
            D.batch = V.n;
            D.batch_first = 0;
            D.first_id = 0;
            D.seed = vs[v].use_model_seed ? (uint64_t)cm.seed : vs[v].seed;
            D.maxt = (double)cm.max_time;
            D.tick = (double)cm.time_tick;
            D.inv_tick = 1.0 / D.tick;
            D.rscale = std::ldexp(1.0, -rt.E);
            D.lay = blob.lay;
            D.variant_id = vs[v].variant_id;
            D.N = cm.N;
            D.K = cm.K;
            D.W = rs.W;
            D.FP = rs.FP;
            D.FT = rs.FT;
            D.sigma = rt.sigma;
            if (C.sh.state_hbm) {
                // one column per thread of the class's largest grid (cus x blocks_per_cu blocks)
                D.lane_state = C.act.p;
                D.lane_stride = (uint32_t)((uint64_t)c->cus * (uint64_t)C.blocks_per_cu * (uint64_t)C.sh.block);
                // element (array, word) of grid thread t at (array * NW + word) * lane_stride + t:
                // every launched thread needs its own column inside the allocation
                if (C.grid * (uint64_t)C.sh.block > (uint64_t)D.lane_stride ||
                    (uint64_t)D.lane_stride * (uint64_t)C.sh.nw * (C.sh.state_hbm == 2 ? 1u : C.sh.wsum_reg == 1 ? 2u : 3u) >
                        (uint64_t)C.act.cap)
                    throw Error(MBG_ERR_ARG, "internal error: a sweep class's lanes do not fit its lane-state buffer");
            }
            if (errors) {
                D.prob2 = B.prob2.p;
                D.th2 = B.th2.p;
                D.err_s = C.err_s.p;
                D.err_v = C.err_v.p;
                D.err_miss = C.err_miss.p;
                D.sq_scale = rs.sq_scale;
                D.sq_scale_th = rs.sq_scale_th;
                D.FQ = rs.FQ;
                D.FQT = rs.FQT;
                D.err_cap = c->err_cap;
                D.err_stride = (uint32_t)((uint64_t)c->cus * (uint64_t)C.blocks_per_cu * (uint64_t)C.sh.block);
            }
            mbg_run_info& I = V.info;
            I = mbg_run_info{};
            I.windows = rs.W;
            I.readouts = cm.K;
            I.prob_exponent = rs.FP;
            I.th_exponent = rs.FT;
            I.rate_exponent = rt.E;
            I.state_words = cm.NW;
            I.discrete = cm.discrete ? 1 : 0;
            I.n_trajectories = V.n;
            I.simulated = V.n;
            I.time_tick = cm.time_tick;
            I.max_time = cm.max_time;
            I.block = C.sh.block;
            I.tables_in_lds = 0;
            I.lds_bytes = (int)C.smem;
            I.blocks_per_cu = C.blocks_per_cu;
            I.state_hbm = C.sh.state_hbm;
            I.error_columns = errors ? 1 : 0;
            I.sq_exponent = errors ? rs.FQ : 0;
            I.th_sq_exponent = errors ? rs.FQT : 0;
            descs.push_back(D);
        }
        C.descs.alloc(descs.size());
        hip_check(hipMemcpyAsync(C.descs.p, descs.data(), descs.size() * sizeof(Desc), hipMemcpyHostToDevice, main),
                  "hipMemcpyAsync");
    }
    if (!jobs.empty()) {
        SS.jobs.alloc(jobs.size());
        hip_check(hipMemcpyAsync(SS.jobs.p, jobs.data(), jobs.size() * sizeof(MargJob), hipMemcpyHostToDevice, main),
                  "hipMemcpyAsync");
        SS.marg.alloc(marg_total);
    }
    {
        // class clocks: start slots at the maximum (atomic min), end slots at 0 (atomic max)
        std::vector<uint64_t> clk0(2 * sw->classes.size());
        for (size_t i = 0; i < clk0.size(); i++) clk0[i] = i % 2 ? 0ull : ~0ull;
        SS.clk.alloc(clk0.size());
        hip_check(hipMemcpyAsync(SS.clk.p, clk0.data(), clk0.size() * 8, hipMemcpyHostToDevice, main), "hipMemcpyAsync");
        hip_check(hipStreamSynchronize(main), "sweep set-up");
    }
    // one stream per class, forked from and joined to the main stream inside the graph
    while (c->class_streams.size() < sw->classes.size()) {
        hipStream_t st;
        hip_check(hipStreamCreate(&st), "hipStreamCreate");
        c->class_streams.push_back(st);
    }
    CaptureScope cap;
    cap.stream = main;
    cap.events.assign(2 * sw->classes.size() + 2, nullptr);
    std::vector<hipEvent_t>& evs = cap.events;
    for (auto& e : evs) hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
    hip_check(hipStreamBeginCapture(main, hipStreamCaptureModeGlobal), "hipStreamBeginCapture");
    cap.capturing = true;
    hip_check(hipEventRecord(evs[0], main), "hipEventRecord");
    // Every class's initial-state pass first, all of them joined before any class kernel
    // starts: a pass launched behind persistent class kernels only gets the CUs they leave,
    // and waited there up to 280 ms.  Then the class kernels, the most work first.
    const size_t ncls = sw->classes.size();
    std::vector<size_t> order(ncls);
    for (size_t ci = 0; ci < ncls; ci++) order[ci] = ci;
    std::stable_sort(order.begin(), order.end(),
                     [&](size_t a, size_t b) { return sw->classes[a]->work > sw->classes[b]->work; });
    for (size_t ci = 0; ci < ncls; ci++) {
        SweepClass& C = *sw->classes[ci];
        hipStream_t st = c->class_streams[ci];
        hip_check(hipStreamWaitEvent(st, evs[0], 0), "hipStreamWaitEvent");
        if (C.init_rates_multi) {
            // initial states of all the class's variants in one launch (64 trajectories per block)
            uint64_t bx = 1;
            for (int v : C.members) bx = std::max<uint64_t>(bx, (sw->vars[v]->D.batch + 63) / 64);
            const mbg_desc_ptr dp = C.descs.p;
            void* iargs[] = {(void*)&dp};
            hip_check(hipModuleLaunchKernel(C.init_rates_multi, (unsigned)bx, (unsigned)C.members.size(), 1, 256, 1, 1, 0,
                                            st, iargs, nullptr),
                      "launch mbg_init_rates_multi");
        } else {
            for (int v : C.members) {
                Desc& D = sw->vars[v]->D;
                const uint64_t nwc = (uint64_t)sw->vars[v]->NWc;
                const uint64_t init_threads = D.batch * nwc;
                launch(C.init_state, (init_threads + 255) / 256, 256, 0, st, D, "launch mbg_init_state");
                launch(C.init_rates, (init_threads + 255) / 256, 256, 0, st, D, "launch mbg_init_rates");
                launch(C.init_readouts, (D.batch + 255) / 256, 256, 0, st, D, "launch mbg_init_readouts");
            }
        }
        hip_check(hipEventRecord(evs[ncls + 1 + ci], st), "hipEventRecord");
        hip_check(hipStreamWaitEvent(main, evs[ncls + 1 + ci], 0), "hipStreamWaitEvent");
    }
    hip_check(hipEventRecord(evs[2 * ncls + 1], main), "hipEventRecord");
    const bool serial = env_int("MBG_TUNE_SWEEP_SERIAL", 0) != 0;  // class kernels one after another
    for (size_t oi = 0; oi < ncls; oi++) {
        const size_t ci = order[oi];
        SweepClass& C = *sw->classes[ci];
        hipStream_t st = serial ? main : c->class_streams[ci];
        if (!serial) hip_check(hipStreamWaitEvent(st, evs[2 * ncls + 1], 0), "hipStreamWaitEvent");
        const uint64_t grid = C.grid;
        const mbg_desc_ptr dp = C.descs.p;
        int n_members = (int)C.members.size();
        int wmax_arg = C.wmax;
        uint64_t* clk = SS.clk.p + 2 * ci;
        int tab_words = (int)(C.sh.tables_lds ? C.tab_bytes / 8 : 0);
        const int* bstart = C.sh.tables_lds ? C.block_start.p : nullptr;
        void* args[] = {(void*)&dp, &n_members, &wmax_arg, (void*)&clk, &tab_words, (void*)&bstart};
        hip_check(hipModuleLaunchKernel(C.sim, (unsigned)grid, 1, 1, (unsigned)C.sh.block, 1, 1, (unsigned)C.smem, st,
                                        args, nullptr),
                  "launch mbg_simulate_multi");
        if (!serial) {
            hip_check(hipEventRecord(evs[ci + 1], st), "hipEventRecord");
            hip_check(hipStreamWaitEvent(main, evs[ci + 1], 0), "hipStreamWaitEvent");
        }
    }
    if (!jobs.empty()) {
        // every variant's readout marginals, reduced from its window table on the device
        const MargJob* jp = SS.jobs.p;
        uint64_t* mp = SS.marg.p;
        void* margs[] = {(void*)&jp, (void*)&mp};
        hip_check(hipModuleLaunchKernel(sw->classes[0]->marginals_batch, (unsigned)w_max, (unsigned)jobs.size(), 1, 256, 1,
                                        1, 0, main, margs, nullptr),
                  "launch mbg_marginals_batch");
    }
    cap.capturing = false;
    hip_check(hipStreamEndCapture(main, &cap.graph), "hipStreamEndCapture");
    hip_check(hipGraphInstantiate(&cap.exec, cap.graph, nullptr, nullptr, 0), "hipGraphInstantiate");
    const auto t1 = std::chrono::steady_clock::now();
    EventPair ev;
    hip_check(hipEventRecord(ev.a, main), "hipEventRecord");
    hip_check(hipGraphLaunch(cap.exec, main), "hipGraphLaunch");
    hip_check(hipEventRecord(ev.b, main), "hipEventRecord");
    hip_check(hipStreamSynchronize(main), "variant sweep");
    float ms = 0.f;
    hip_check(hipEventElapsedTime(&ms, ev.a, ev.b), "hipEventElapsedTime");

    mbg_sweep_info& SI = sw->info;
    SI = mbg_sweep_info{};
    SI.variants = nv;
    SI.classes = (int)sw->classes.size();
    SI.graph_ms = ms;
    SI.setup_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    std::vector<uint64_t> counters((size_t)2 * nv), marg(marg_total), clk(2 * sw->classes.size());
    hip_check(hipMemcpy(counters.data(), SS.counters.p, counters.size() * 8, hipMemcpyDeviceToHost), "hipMemcpy");
    hip_check(hipMemcpy(clk.data(), SS.clk.p, clk.size() * 8, hipMemcpyDeviceToHost), "hipMemcpy");
    for (size_t ci = 0; ci < sw->classes.size(); ci++) {
        SweepClass& C = *sw->classes[ci];
        C.kernel_ms = clk[2 * ci + 1] > clk[2 * ci] && c->wall_khz > 0
                          ? (double)(clk[2 * ci + 1] - clk[2 * ci]) / (double)c->wall_khz
                          : 0.0;
        C.trajectories = C.transitions = 0;
        for (int v : C.members) {
            C.trajectories += sw->vars[v]->n;
            C.transitions += counters[2 * (size_t)v + 1];
        }
    }
    if (marg_total) hip_check(hipMemcpy(marg.data(), SS.marg.p, marg_total * 8, hipMemcpyDeviceToHost), "hipMemcpy");
    for (int v = 0; v < nv; v++) {
        VariantRun& V = *sw->vars[v];
        const uint64_t nfp = counters[2 * (size_t)v], flips = counters[2 * (size_t)v + 1];
        V.marg_raw.clear();
        if (job_of[v] >= 0) {
            const MargJob& J = jobs[(size_t)job_of[v]];
            V.marg_raw.assign(marg.begin() + (ptrdiff_t)J.out_off,
                              marg.begin() + (ptrdiff_t)(J.out_off + (uint64_t)J.W * (uint64_t)J.K));
        }
        if (nfp > V.D.fp_cap) throw Error(MBG_ERR_LIMIT, "fixed-point buffer overflow");
        collect_fixed_points(c, *V.mod, V.buf.fp.p, nfp, V.NWc, V.m->cm.NW, main, V.fixed_states, V.fixed_counts);
        V.info.transitions = flips;
        c->variant_cost[variant_key(vs[v])] = (double)flips / (double)V.n;
        V.info.fixed_point_trajectories = nfp;
        V.info.kernel_ms = ms;
        SI.trajectories += V.n;
        SI.transitions += flips;
    }
    c->sweep = std::move(sw);
    if (errors) {
        // a lane's window list overflowed: repeat the sweep with lists four times longer (the
        // sweep is deterministic, so the repeat gives the same tables)
        uint64_t miss = 0;
        for (auto& Cp : c->sweep->classes) {
            uint64_t m = 0;
            hip_check(hipMemcpy(&m, Cp->err_miss.p, 8, hipMemcpyDeviceToHost), "hipMemcpy");
            miss += m;
        }
        if (miss) {
            if (c->err_cap >= 4096 || depth > 6)
                throw Error(MBG_ERR_LIMIT, "error columns: a trajectory visits more than 4,096 projected states in one window");
            c->err_cap *= 4;
            run_variants_impl(c, vs, nv, flags, depth + 1);
        }
    }
}

// Sweep autotuner: for each word class of the sweep, the class's variants alone (their full
// trajectory counts) under candidate layouts, by coordinate descent from the default rule: the
// tables in LDS or HBM, the active sets in HBM or LDS, the dependents cap (none, 3 or the
// default).  A
// candidate replaces the best when its class kernel time is more than 1 % shorter.  The
// choices are kept per class key and applied by later sweeps that form the same class.
void autotune_variants_impl(mbg_ctx* c, const mbg_variant* vs, int nv, int flags) {
    std::map<std::pair<int, int>, std::vector<mbg_variant>> groups;
    for (int v = 0; v < nv; v++) {
        if (!vs[v].model) throw Error(MBG_ERR_ARG, "variant without a model");
        const mbg::CompiledModel& cm = vs[v].model->cm;
        groups[{word_class(cm.NW), cm.discrete ? 1 : 0}].push_back(vs[v]);
    }
    struct Reset {
        mbg_ctx* c;
        ~Reset() { c->force_class = nullptr; }
    } reset{c};
    for (auto& g : groups) {
        std::vector<mbg_variant>& sub = g.second;
        auto time_of = [&](const ClassChoice& cand, std::vector<long long>* key, Shape* shape) {
            c->force_class = &cand;
            run_variants_impl(c, sub.data(), (int)sub.size(), flags);
            c->force_class = nullptr;
            const SweepClass& C = *c->sweep->classes[0];
            if (key) *key = C.key;
            if (shape) *shape = C.sh;
            return C.kernel_ms > 0 ? C.kernel_ms : (double)c->sweep->info.graph_ms;
        };
        ClassChoice best;
        std::vector<long long> key;
        Shape sh{};
        time_of(best, nullptr, nullptr);  // compile and warm the default layout
        double best_ms = time_of(best, &key, &sh);
        // resolve the default rule into explicit choices, then flip one decision at a time
        best.tables_lds = sh.tables_lds;
        best.state_hbm = sh.state_hbm == 2 ? 2 : 0;
        best.dep_cap = sh.dep_cap;
        auto consider = [&](ClassChoice cand) {
            Shape got{};
            time_of(cand, nullptr, nullptr);  // compile
            const double ms = time_of(cand, nullptr, &got);
            // a flip the layout rules refused (the tables or active sets do not fit) is no candidate
            if (got.tables_lds != cand.tables_lds || (cand.state_hbm == 2) != (got.state_hbm == 2)) return;
            if (ms < best_ms * 0.99) {
                best_ms = ms;
                best = cand;
            }
        };
        ClassChoice t = best;
        t.tables_lds = !best.tables_lds;
        consider(t);
        t = best;
        t.state_hbm = best.state_hbm == 2 ? 0 : 2;
        consider(t);
        for (int cap : {0, 3}) {
            if (cap == best.dep_cap) continue;
            t = best;
            t.dep_cap = cap;
            consider(t);
        }
        c->tuned_classes[key] = best;
    }
}

const VariantRun& sweep_variant(const mbg_ctx* c, int v) {
    if (!c->sweep) throw Error(MBG_ERR_NOT_BUILT, "no variant sweep has run");
    if (v < 0 || v >= (int)c->sweep->vars.size()) throw Error(MBG_ERR_ARG, "variant index out of range");
    return *c->sweep->vars[v];
}

// Host copies of one sweep variant's tables, in the form the output writers read.
struct VariantHost {
    const VariantRun& V;
    mbg::PinnedU64 prob;
    std::vector<uint64_t> th, fin, prob2, th2;

    VariantHost(const mbg_ctx* c, const VariantRun& v) : V(v) {
        hip_check(hipSetDevice(c->device), "hipSetDevice");
        const size_t W = (size_t)V.info.windows, S = (size_t)1 << V.info.readouts;
        prob.resize(W * S);
        th.resize(W);
        fin.resize(S);
        hip_check(hipMemcpy(prob.data(), V.buf.prob.p, W * S * 8, hipMemcpyDeviceToHost), "hipMemcpy");
        hip_check(hipMemcpy(th.data(), V.buf.th.p, W * 8, hipMemcpyDeviceToHost), "hipMemcpy");
        hip_check(hipMemcpy(fin.data(), V.buf.fin.p, S * 8, hipMemcpyDeviceToHost), "hipMemcpy");
        if (V.info.error_columns) {
            prob2.resize(W * S);
            th2.resize(W);
            hip_check(hipMemcpy(prob2.data(), V.buf.prob2.p, W * S * 8, hipMemcpyDeviceToHost), "hipMemcpy");
            hip_check(hipMemcpy(th2.data(), V.buf.th2.p, W * 8, hipMemcpyDeviceToHost), "hipMemcpy");
        }
    }
    mbg::Results view() const {
        mbg::Results r;
        r.info = &V.info;
        r.prob = &prob;
        r.th = &th;
        r.fin = &fin;
        if (V.info.error_columns) {
            r.prob2 = &prob2;
            r.th2 = &th2;
        }
        r.fixed_states = &V.fixed_states;
        r.fixed_counts = &V.fixed_counts;
        return r;
    }
};

void normalise_marginals(const mbg_run_info& I, const std::vector<uint64_t>& raw, double* out) {
    for (size_t i = 0; i < raw.size(); i++) {
        const double v = (double)raw[i];
        out[i] = I.discrete ? v / (double)I.n_trajectories
                            : std::ldexp(v, -I.prob_exponent) / ((double)I.n_trajectories * (double)I.time_tick);
    }
}

// The window table of the last run on the host: copied from the device on first use (a run
// leaves it in device memory, so callers that only need device-side results or the small
// tables never pay for the transfer).
void ensure_prob(const mbg_ctx* cc) {
    mbg_ctx* c = const_cast<mbg_ctx*>(cc);
    if (c->prob_on_host) return;
    hip_check(hipSetDevice(c->device), "hipSetDevice");
    if (!c->prob.empty())
        hip_check(hipMemcpy(c->prob.data(), c->buf.prob.p, c->prob.size() * 8, hipMemcpyDeviceToHost), "hipMemcpy");
    c->prob_on_host = true;
}

mbg::Results results_view(const mbg_ctx* c) {
    ensure_prob(c);
    mbg::Results r;
    r.info = &c->info;
    r.prob = &c->prob;
    r.th = &c->th;
    r.fin = &c->fin;
    if (c->sparse) {
        r.cell_w = &c->cell_w;
        r.cell_s = &c->cell_s;
        r.cell_v = &c->cell_v;
        r.fin_s = &c->fin_s;
        r.fin_v = &c->fin_v;
    }
    if (c->info.error_columns) {
        r.th2 = &c->th2;
        if (c->sparse) r.cell_v2 = &c->cell_v2;
        else r.prob2 = &c->prob2;
    }
    r.fixed_states = &c->fixed_states;
    r.fixed_counts = &c->fixed_counts;
    return r;
}

std::string read_file(const char* path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw Error(MBG_ERR_ARG, std::string("cannot open ") + path);
    std::stringstream ss;
    ss << f.rdbuf();
    return ss.str();
}

template <class F>
mbg_status guarded(mbg_ctx* c, F&& f) {
    try {
        f();
        if (c) c->err.clear();
        return MBG_OK;
    } catch (const Error& e) {
        if (c) c->err = e.what();
        return e.code;
    } catch (const mbg::ParseError& e) {
        if (c) c->err = e.what();
        return MBG_ERR_PARSE;
    } catch (const std::bad_alloc&) {
        if (c) c->err = "out of host memory";
        return MBG_ERR_HIP;
    } catch (const std::exception& e) {
        if (c) c->err = e.what();
        return MBG_ERR_ARG;
    }
}

mbg_status parse_impl(const std::string& bnd, const std::string& cfg, mbg_model** out, char* err, size_t errlen) {
    *out = nullptr;
    auto m = std::make_unique<mbg_model>();
    try {
        m->parsed = std::make_shared<mbg::ParsedModel>();
        mbg::parse_model(bnd, cfg, *m->parsed);
        m->cm = mbg::compile_model(*m->parsed);
    } catch (const mbg::ParseError& e) {
        copy_err(err, errlen, e.what());
        return MBG_ERR_PARSE;
    } catch (const std::exception& e) {
        copy_err(err, errlen, e.what());
        return MBG_ERR_PARSE;
    }
    if (m->cm.K > mbg::kMaxSparseReadouts) {
        copy_err(err, errlen, mbg::kReadoutLimitMessage);
        return MBG_ERR_LIMIT;
    }
    *out = m.release();
    return MBG_OK;
}

}  // namespace

extern "C" {

mbg_status mbg_model_parse(const char* bnd_text, const char* cfg_text, mbg_model** out, char* err, size_t errlen) {
    if (!bnd_text || !cfg_text || !out) return MBG_ERR_ARG;
    return parse_impl(bnd_text, cfg_text, out, err, errlen);
}

mbg_status mbg_model_parse_files(const char* bnd_path, const char* cfg_path, mbg_model** out, char* err,
                                 size_t errlen) {
    if (!bnd_path || !cfg_path || !out) return MBG_ERR_ARG;
    std::string b, c;
    try {
        b = read_file(bnd_path);
        c = read_file(cfg_path);
    } catch (const std::exception& e) {
        *out = nullptr;
        copy_err(err, errlen, e.what());
        return MBG_ERR_ARG;
    }
    return parse_impl(b, c, out, err, errlen);
}

void mbg_model_free(mbg_model* m) { delete m; }

mbg_status mbg_variants_parse(const mbg_model* base, const char* text, size_t len, mbg_variant* out, char* names,
                              int cap, int* count, char* err, size_t errlen) {
    if (!base || !text || !count || cap < 0) return MBG_ERR_ARG;
    *count = 0;
    std::vector<mbg::VariantSpec> specs;
    try {
        specs = mbg::parse_variants(std::string(text, len), base->cm.names);
    } catch (const std::exception& e) {
        copy_err(err, errlen, std::string("variants file, ") + e.what());
        return MBG_ERR_PARSE;
    }
    *count = (int)specs.size();
    if (!out) return MBG_OK;
    if ((int)specs.size() > cap) {
        copy_err(err, errlen, "variants file holds " + std::to_string(specs.size()) + " variants, room for " +
                                  std::to_string(cap));
        return MBG_ERR_LIMIT;
    }
    const int N = base->cm.N;
    std::vector<mbg_model*> made;
    for (const mbg::VariantSpec& s : specs) {
        auto v = std::make_unique<mbg_model>(*base);
        std::vector<float> up(N), down(N), ist(N);
        for (int i = 0; i < N; i++) {
            up[i] = (base->mult_up.empty() ? 1.f : base->mult_up[i]) * s.mult_up[i];
            down[i] = (base->mult_down.empty() ? 1.f : base->mult_down[i]) * s.mult_down[i];
            ist[i] = !std::isnan(s.istate[i]) ? s.istate[i] : base->istate.empty() ? base->cm.istate[i] : base->istate[i];
        }
        if (mbg_model_set_variant(v.get(), up.data(), down.data(), ist.data()) != MBG_OK) {
            for (mbg_model* p : made) delete p;
            copy_err(err, errlen, "variants file, line " + std::to_string(s.line) + ": variant '" + s.name +
                                      "' gives rates out of range");
            return MBG_ERR_PARSE;
        }
        made.push_back(v.release());
    }
    for (size_t i = 0; i < specs.size(); i++) {
        const mbg::VariantSpec& s = specs[i];
        out[i].model = made[i];
        out[i].n_trajectories = s.has_count ? s.count : 0;
        out[i].seed = s.seed;
        out[i].use_model_seed = s.has_seed ? 0 : 1;
        out[i].variant_id = 0;
        if (names) {
            char* dst = names + 64 * i;
            const size_t k = std::min<size_t>(s.name.size(), 63);
            std::memcpy(dst, s.name.data(), k);
            dst[k] = '\0';
        }
    }
    return MBG_OK;
}

mbg_status mbg_variants_parse_file(const mbg_model* base, const char* path, mbg_variant* out, char* names, int cap,
                                   int* count, char* err, size_t errlen) {
    if (!base || !path || !count) return MBG_ERR_ARG;
    std::string text;
    try {
        text = read_file(path);
    } catch (const std::exception& e) {
        *count = 0;
        copy_err(err, errlen, e.what());
        return MBG_ERR_ARG;
    }
    return mbg_variants_parse(base, text.data(), text.size(), out, names, cap, count, err, errlen);
}

mbg_status mbg_model_clone(const mbg_model* m, mbg_model** out) {
    if (!m || !out) return MBG_ERR_ARG;
    try {
        auto c = std::make_unique<mbg_model>();
        c->parsed = m->parsed;
        c->cm = m->cm;
        c->mult_up = m->mult_up;
        c->mult_down = m->mult_down;
        c->istate = m->istate;
        *out = c.release();
    } catch (const std::bad_alloc&) {
        *out = nullptr;
        return MBG_ERR_HIP;
    }
    return MBG_OK;
}
int mbg_model_num_nodes(const mbg_model* m) { return m ? m->cm.N : 0; }
const char* mbg_model_node_name(const mbg_model* m, int i) {
    return (m && i >= 0 && i < m->cm.N) ? m->cm.names[i].c_str() : nullptr;
}
float mbg_model_node_istate(const mbg_model* m, int i) {
    return (m && i >= 0 && i < m->cm.N) ? m->cm.istate[i] : NAN;
}
int mbg_model_node_is_internal(const mbg_model* m, int i) {
    return (m && i >= 0 && i < m->cm.N) ? m->cm.internal[i] : -1;
}
int mbg_model_node_is_generic(const mbg_model* m, int i) {
    return (m && i >= 0 && i < m->cm.N) ? m->cm.generic[i] : -1;
}
int mbg_model_num_readouts(const mbg_model* m) { return m ? m->cm.K : 0; }

mbg_status mbg_model_constant(const mbg_model* m, const char* name, double* out) {
    if (!m || !name || !out) return MBG_ERR_ARG;
    auto it = m->parsed->constants.find(name);
    if (it == m->parsed->constants.end()) return MBG_ERR_ARG;
    *out = it->second;
    return MBG_OK;
}

mbg_status mbg_model_set_variant(mbg_model* m, const float* mult_up, const float* mult_down, const float* istate) {
    if (!m) return MBG_ERR_ARG;
    const int N = m->cm.N;
    auto put = [N](std::vector<float>& v, const float* p) {
        if (p) v.assign(p, p + N);
        else v.clear();
    };
    put(m->mult_up, mult_up);
    put(m->mult_down, mult_down);
    put(m->istate, istate);
    try {
        (void)mbg::make_rate_tables(m->cm, m->mult_up.empty() ? nullptr : m->mult_up.data(),
                                    m->mult_down.empty() ? nullptr : m->mult_down.data(), nullptr);
    } catch (const std::exception&) {
        m->mult_up.clear();
        m->mult_down.clear();
        return MBG_ERR_ARG;
    }
    return MBG_OK;
}

mbg_status mbg_model_get_variant(const mbg_model* m, float* mult_up, float* mult_down, float* istate) {
    if (!m) return MBG_ERR_ARG;
    for (int i = 0; i < m->cm.N; i++) {
        if (mult_up) mult_up[i] = m->mult_up.empty() ? 1.f : m->mult_up[i];
        if (mult_down) mult_down[i] = m->mult_down.empty() ? 1.f : m->mult_down[i];
        if (istate) istate[i] = m->istate.empty() ? m->cm.istate[i] : m->istate[i];
    }
    return MBG_OK;
}

mbg_status mbg_model_transition_rates(const mbg_model* m, const uint64_t* state, float* rates) {
    if (!m || !state || !rates) return MBG_ERR_ARG;
    for (int i = 0; i < m->cm.N; i++) rates[i] = mbg::node_rate(m->cm, i, state);
    return MBG_OK;
}

mbg_status mbg_model_generate_source(const mbg_model* m, int trace, char* buf, size_t cap, size_t* len) {
    if (!m) return MBG_ERR_ARG;
    const mbg::CompiledModel& cm = m->cm;
    Shape sh{cm.NW, 64, cm.discrete ? 1 : 0, trace ? 1 : 0, cm.any_generic ? 1 : 0, 0, std::max(1, cm.max_inputs)};
    sh.sparse = cm.K > mbg::kMaxReadouts ? 1 : 0;
    sh.prog = cm.any_prog ? 1 : 0;
    if (env_int("MBG_TUNE_GENCODE", 0) != 0 && m->parsed) sh.gen = generate_logic_source(*m->parsed, cm);
    const std::string s = generate_source(cm, sh);
    if (len) *len = s.size();
    if (buf && cap) {
        const size_t k = std::min(cap - 1, s.size());
        std::memcpy(buf, s.data(), k);
        buf[k] = 0;
    }
    return MBG_OK;
}

mbg_status mbg_create(int device, mbg_ctx** out) {
    if (!out) return MBG_ERR_ARG;
    *out = nullptr;
    auto c = std::make_unique<mbg_ctx>();
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return MBG_ERR_HIP;
    if (device < 0 || device >= n) return MBG_ERR_ARG;
    c->device = device;
    if (hipSetDevice(device) != hipSuccess) return MBG_ERR_HIP;
    if (hipDeviceGetAttribute(&c->cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) return MBG_ERR_HIP;
    if (hipDeviceGetAttribute(&c->lds_per_block, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess)
        return MBG_ERR_HIP;
    if (hipDeviceGetAttribute(&c->wall_khz, hipDeviceAttributeWallClockRate, device) != hipSuccess) c->wall_khz = 0;
    if (hipStreamCreate(&c->stream) != hipSuccess) return MBG_ERR_HIP;
    *out = c.release();
    return MBG_OK;
}

void mbg_destroy(mbg_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    for (auto& kv : c->modules) (void)hipModuleUnload(kv.second.mod);
    c->sweep.reset();
    for (hipStream_t s : c->class_streams) (void)hipStreamDestroy(s);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char* mbg_last_error(const mbg_ctx* c) { return c ? c->err.c_str() : "null context"; }

mbg_status mbg_compile(mbg_ctx* c, const mbg_model* m, int trace) {
    if (!c || !m) return MBG_ERR_ARG;
    return guarded(c, [&] {
        hip_check(hipSetDevice(c->device), "hipSetDevice");
        const mbg::CompiledModel& cm = m->cm;
        const mbg::RateTables rt = mbg::make_rate_tables(cm, m->mult_up.empty() ? nullptr : m->mult_up.data(),
                                                         m->mult_down.empty() ? nullptr : m->mult_down.data(),
                                                         m->istate.empty() ? nullptr : m->istate.data());
        const auto tu = c->tuned.find(tune_key(cm, rt, trace));
        get_module(c, cm, tu != c->tuned.end() ? tu->second : default_shape(c, cm, rt, 0, trace));
    });
}

mbg_status mbg_run(mbg_ctx* c, const mbg_model* m, const mbg_run_params* p) {
    if (!c || !m || !p) return MBG_ERR_ARG;
    return guarded(c, [&] {
        hip_check(hipSetDevice(c->device), "hipSetDevice");
        c->have = false;
        run_impl(c, m, p);
    });
}

mbg_status mbg_autotune(mbg_ctx* c, const mbg_model* m, uint64_t sample, int trace, mbg_tune_info* out) {
    if (!c || !m) return MBG_ERR_ARG;
    return guarded(c, [&] {
        hip_check(hipSetDevice(c->device), "hipSetDevice");
        c->have = false;
        autotune_impl(c, m, sample, trace, out);
    });
}

mbg_status mbg_get_run_info(const mbg_ctx* c, mbg_run_info* out) {
    if (!c || !out) return MBG_ERR_ARG;
    if (!c->have) return MBG_ERR_NOT_BUILT;
    *out = c->info;
    return MBG_OK;
}

mbg_status mbg_get_probtraj_raw(const mbg_ctx* c, uint64_t* prob, uint64_t* th) {
    if (!c) return MBG_ERR_ARG;
    if (!c->have) return MBG_ERR_NOT_BUILT;
    return guarded(const_cast<mbg_ctx*>(c), [&] {
        if (prob && c->sparse) throw Error(MBG_ERR_LIMIT, kSparseMessage);
        if (prob) {
            ensure_prob(c);
            std::memcpy(prob, c->prob.data(), c->prob.size() * 8);
        }
        if (th) std::memcpy(th, c->th.data(), c->th.size() * 8);
    });
}

mbg_status mbg_get_probtraj(const mbg_ctx* c, double* prob, double* th) {
    if (!c) return MBG_ERR_ARG;
    if (!c->have) return MBG_ERR_NOT_BUILT;
    return guarded(const_cast<mbg_ctx*>(c), [&] {
        if (prob && c->sparse) throw Error(MBG_ERR_LIMIT, kSparseMessage);
        const mbg::Results r = results_view(c);
        if (prob)
            for (size_t i = 0; i < c->prob.size(); i++) prob[i] = mbg::window_prob(r, i);
        if (th)
            for (size_t w = 0; w < c->th.size(); w++) th[w] = mbg::window_th(r, w);
    });
}

mbg_status mbg_get_probtraj_errors(const mbg_ctx* c, double* err_prob, double* err_th) {
    if (!c) return MBG_ERR_ARG;
    if (!c->have || !c->info.error_columns) return MBG_ERR_NOT_BUILT;
    return guarded(const_cast<mbg_ctx*>(c), [&] {
        if (err_prob && c->sparse) throw Error(MBG_ERR_LIMIT, kSparseMessage);
        const mbg::Results r = results_view(c);
        if (err_prob)
            for (size_t i = 0; i < c->prob.size(); i++) err_prob[i] = mbg::window_prob_error(r, i);
        if (err_th)
            for (size_t w = 0; w < c->th.size(); w++) err_th[w] = mbg::window_th_error(r, w);
    });
}

mbg_status mbg_get_probtraj_squares_raw(const mbg_ctx* c, uint64_t* prob_sq, uint64_t* th_sq) {
    if (!c) return MBG_ERR_ARG;
    if (!c->have || !c->info.error_columns) return MBG_ERR_NOT_BUILT;
    if (prob_sq && c->sparse) {
        const_cast<mbg_ctx*>(c)->err = kSparseMessage;
        return MBG_ERR_LIMIT;
    }
    if (prob_sq) std::memcpy(prob_sq, c->prob2.data(), c->prob2.size() * 8);
    if (th_sq) std::memcpy(th_sq, c->th2.data(), c->th2.size() * 8);
    return MBG_OK;
}

mbg_status mbg_get_probtraj_cell_squares(const mbg_ctx* c, uint64_t* squares, size_t cap, size_t* n) {
    if (!c) return MBG_ERR_ARG;
    if (!c->have || !c->info.error_columns) return MBG_ERR_NOT_BUILT;
    size_t k = 0;
    auto put = [&](uint64_t q) {
        if (k < cap && squares) squares[k] = q;
        k++;
    };
    if (c->sparse) {
        for (uint64_t q : c->cell_v2) put(q);
    } else {
        ensure_prob(c);
        for (size_t i = 0; i < c->prob.size(); i++)
            if (c->prob[i]) put(c->prob2[i]);
    }
    if (n) *n = k;
    return MBG_OK;
}

mbg_status mbg_get_final_states(const mbg_ctx* c, uint64_t* counts) {
    if (!c || !counts) return MBG_ERR_ARG;
    if (!c->have) return MBG_ERR_NOT_BUILT;
    if (c->sparse) {
        const_cast<mbg_ctx*>(c)->err = kSparseMessage;
        return MBG_ERR_LIMIT;
    }
    std::memcpy(counts, c->fin.data(), c->fin.size() * 8);
    return MBG_OK;
}

mbg_status mbg_get_probtraj_cells(const mbg_ctx* c, uint32_t* windows, uint64_t* states, uint64_t* cells, size_t cap,
                                  size_t* n) {
    if (!c) return MBG_ERR_ARG;
    if (!c->have) return MBG_ERR_NOT_BUILT;
    return guarded(const_cast<mbg_ctx*>(c), [&] {
        size_t k = 0;
        auto put = [&](uint32_t w, uint64_t st, uint64_t v) {
            if (k < cap) {
                if (windows) windows[k] = w;
                if (states) states[k] = st;
                if (cells) cells[k] = v;
            }
            k++;
        };
        if (c->sparse) {
            for (size_t i = 0; i < c->cell_v.size(); i++) put(c->cell_w[i], c->cell_s[i], c->cell_v[i]);
        } else {
            ensure_prob(c);
            const uint64_t S = 1ull << c->info.readouts;
            for (size_t i = 0; i < c->prob.size(); i++)
                if (c->prob[i]) put((uint32_t)(i / S), i % S, c->prob[i]);
        }
        if (n) *n = k;
    });
}

mbg_status mbg_get_final_cells(const mbg_ctx* c, uint64_t* states, uint64_t* counts, size_t cap, size_t* n) {
    if (!c) return MBG_ERR_ARG;
    if (!c->have) return MBG_ERR_NOT_BUILT;
    size_t k = 0;
    auto put = [&](uint64_t st, uint64_t v) {
        if (k < cap) {
            if (states) states[k] = st;
            if (counts) counts[k] = v;
        }
        k++;
    };
    if (c->sparse) {
        for (size_t i = 0; i < c->fin_v.size(); i++) put(c->fin_s[i], c->fin_v[i]);
    } else {
        for (size_t s = 0; s < c->fin.size(); s++)
            if (c->fin[s]) put(s, c->fin[s]);
    }
    if (n) *n = k;
    return MBG_OK;
}

mbg_status mbg_get_fixed_points(const mbg_ctx* c, uint64_t* states, uint64_t* counts, size_t cap, size_t* n) {
    if (!c) return MBG_ERR_ARG;
    if (!c->have) return MBG_ERR_NOT_BUILT;
    const size_t F = c->fixed_counts.size(), NW = (size_t)c->info.state_words;
    if (n) *n = F;
    const size_t k = std::min(cap, F);
    if (states && k) std::memcpy(states, c->fixed_states.data(), k * NW * 8);
    if (counts && k) std::memcpy(counts, c->fixed_counts.data(), k * 8);
    return MBG_OK;
}

mbg_status mbg_get_summaries(const mbg_ctx* c, uint64_t* out, size_t count) {
    if (!c || !out) return MBG_ERR_ARG;
    if (!c->have) return MBG_ERR_NOT_BUILT;
    if (count * 4 > c->summ.size()) return MBG_ERR_ARG;
    std::memcpy(out, c->summ.data(), count * 4 * 8);
    return MBG_OK;
}

mbg_status mbg_write_csv(const mbg_ctx* c, const mbg_model* m, const char* prefix) {
    if (!c || !m || !prefix) return MBG_ERR_ARG;
    if (!c->have) return MBG_ERR_NOT_BUILT;
    return guarded(const_cast<mbg_ctx*>(c), [&] {
        const mbg::Results r = results_view(c);
        if (!mbg::write_probtraj_csv(r, m->cm, std::string(prefix) + "_probtraj.csv") ||
            !mbg::write_fp_csv(r, m->cm, std::string(prefix) + "_fp.csv"))
            throw Error(MBG_ERR_ARG, std::string("cannot write output files with prefix ") + prefix);
    });
}

mbg_status mbg_write_text(const mbg_ctx* c, const mbg_model* m, const char* path) {
    if (!c || !m || !path) return MBG_ERR_ARG;
    if (!c->have) return MBG_ERR_NOT_BUILT;
    return guarded(const_cast<mbg_ctx*>(c), [&] {
        const mbg::Results r = results_view(c);
        if (std::string(path) == "-") {
            mbg::write_text(r, m->cm, std::cout);
        } else {
            std::ofstream f(path);
            if (!f) throw Error(MBG_ERR_ARG, std::string("cannot open ") + path);
            mbg::write_text(r, m->cm, f);
        }
    });
}

}  // extern "C"

extern "C" {

mbg_status mbg_get_marginals(const mbg_ctx* c, double* out) {
    if (!c || !out) return MBG_ERR_ARG;
    if (!c->have) return MBG_ERR_NOT_BUILT;
    return guarded(const_cast<mbg_ctx*>(c), [&] {
        mbg_ctx* cc = const_cast<mbg_ctx*>(c);
        const int W = c->info.windows, K = c->info.readouts;
        std::vector<uint64_t> raw((size_t)W * K, 0);
        if (c->sparse) {
            for (size_t i = 0; i < c->cell_v.size(); i++)
                for (int j = 0; j < K; j++)
                    if ((c->cell_s[i] >> j) & 1ull) raw[(size_t)c->cell_w[i] * K + j] += c->cell_v[i];
        } else if (K > 0) {
            // reduced on the device from the window table (no transfer of the whole table)
            hip_check(hipSetDevice(c->device), "hipSetDevice");
            cc->marg.alloc(raw.size());
            const void* prob = cc->buf.prob.p;
            void* outp = cc->marg.p;
            int k = K;
            void* args[] = {(void*)&prob, &k, &outp};
            hip_check(hipModuleLaunchKernel(c->marg_fn, (unsigned)W, 1, 1, 256, 1, 1, 0, c->stream, args, nullptr),
                      "launch mbg_marginals");
            hip_check(hipMemcpyAsync(raw.data(), cc->marg.p, raw.size() * 8, hipMemcpyDeviceToHost, c->stream),
                      "hipMemcpyAsync");
            hip_check(hipStreamSynchronize(c->stream), "mbg_marginals");
        }
        normalise_marginals(c->info, raw, out);
    });
}

mbg_status mbg_run_variants(mbg_ctx* c, const mbg_variant* variants, int count) {
    return mbg_run_variants_ex(c, variants, count, 0);
}

mbg_status mbg_run_variants_ex(mbg_ctx* c, const mbg_variant* variants, int count, int flags) {
    if (!c || !variants || count <= 0 || (flags & ~MBG_SWEEP_ERROR_COLUMNS)) return MBG_ERR_ARG;
    return guarded(c, [&] {
        hip_check(hipSetDevice(c->device), "hipSetDevice");
        run_variants_impl(c, variants, count, flags);
    });
}

mbg_status mbg_autotune_variants(mbg_ctx* c, const mbg_variant* variants, int count, int flags) {
    if (!c || !variants || count <= 0) return MBG_ERR_ARG;
    return guarded(c, [&] { autotune_variants_impl(c, variants, count, flags); });
}

mbg_status mbg_get_sweep_info(const mbg_ctx* c, mbg_sweep_info* out) {
    if (!c || !out) return MBG_ERR_ARG;
    if (!c->sweep) return MBG_ERR_NOT_BUILT;
    *out = c->sweep->info;
    return MBG_OK;
}

mbg_status mbg_get_sweep_class_info(const mbg_ctx* c, int i, mbg_sweep_class_info* out) {
    if (!c || !out) return MBG_ERR_ARG;
    if (!c->sweep) return MBG_ERR_NOT_BUILT;
    if (i < 0 || i >= (int)c->sweep->classes.size()) return MBG_ERR_ARG;
    const SweepClass& C = *c->sweep->classes[(size_t)i];
    mbg_sweep_class_info& o = *out;
    o = mbg_sweep_class_info{};
    o.words = C.sh.nw;
    o.discrete = C.sh.discrete;
    o.variants = (int)C.members.size();
    o.trajectories = C.trajectories;
    o.transitions = C.transitions;
    o.kernel_ms = C.kernel_ms;
    o.block = C.sh.block;
    o.blocks_per_cu = C.blocks_per_cu;
    o.state_hbm = C.sh.state_hbm;
    o.wsum_reg = C.sh.wsum_reg;
    o.sumw = C.sh.sumw;
    o.dep_batch = C.sh.dep_batch;
    o.dep_cap = C.sh.dep_cap;
    o.scan = C.sh.scan;
    o.th_lds = C.sh.th_lds;
    o.tuned = C.tuned ? 1 : 0;
    o.tables_lds = C.sh.tables_lds;
    return MBG_OK;
}

mbg_status mbg_get_variant_info(const mbg_ctx* c, int v, mbg_run_info* out) {
    if (!c || !out) return MBG_ERR_ARG;
    return guarded(const_cast<mbg_ctx*>(c), [&] { *out = sweep_variant(c, v).info; });
}

mbg_status mbg_get_variant_probtraj_raw(const mbg_ctx* c, int v, uint64_t* prob, uint64_t* th) {
    if (!c) return MBG_ERR_ARG;
    return guarded(const_cast<mbg_ctx*>(c), [&] {
        const VariantRun& V = sweep_variant(c, v);
        const size_t W = (size_t)V.info.windows, S = (size_t)1 << V.info.readouts;
        if (prob) hip_check(hipMemcpy(prob, V.buf.prob.p, W * S * 8, hipMemcpyDeviceToHost), "hipMemcpy");
        if (th) hip_check(hipMemcpy(th, V.buf.th.p, W * 8, hipMemcpyDeviceToHost), "hipMemcpy");
    });
}

mbg_status mbg_get_variant_probtraj_squares_raw(const mbg_ctx* c, int v, uint64_t* prob_sq, uint64_t* th_sq) {
    if (!c) return MBG_ERR_ARG;
    return guarded(const_cast<mbg_ctx*>(c), [&] {
        const VariantRun& V = sweep_variant(c, v);
        if (!V.info.error_columns) throw Error(MBG_ERR_NOT_BUILT, "the sweep did not accumulate the error columns");
        const size_t W = (size_t)V.info.windows, S = (size_t)1 << V.info.readouts;
        if (prob_sq) hip_check(hipMemcpy(prob_sq, V.buf.prob2.p, W * S * 8, hipMemcpyDeviceToHost), "hipMemcpy");
        if (th_sq) hip_check(hipMemcpy(th_sq, V.buf.th2.p, W * 8, hipMemcpyDeviceToHost), "hipMemcpy");
    });
}

mbg_status mbg_get_variant_marginals(const mbg_ctx* c, int v, double* out) {
    if (!c || !out) return MBG_ERR_ARG;
    return guarded(const_cast<mbg_ctx*>(c), [&] {
        const VariantRun& V = sweep_variant(c, v);
        std::vector<uint64_t> raw = V.marg_raw;
        raw.resize((size_t)V.info.windows * (size_t)V.info.readouts, 0);
        normalise_marginals(V.info, raw, out);
    });
}

mbg_status mbg_get_variant_final_states(const mbg_ctx* c, int v, uint64_t* counts) {
    if (!c || !counts) return MBG_ERR_ARG;
    return guarded(const_cast<mbg_ctx*>(c), [&] {
        const VariantRun& V = sweep_variant(c, v);
        const size_t S = (size_t)1 << V.info.readouts;
        hip_check(hipMemcpy(counts, V.buf.fin.p, S * 8, hipMemcpyDeviceToHost), "hipMemcpy");
    });
}

mbg_status mbg_write_variant_csv(const mbg_ctx* c, int v, const char* prefix) {
    if (!c || !prefix) return MBG_ERR_ARG;
    return guarded(const_cast<mbg_ctx*>(c), [&] {
        const VariantRun& V = sweep_variant(c, v);
        const VariantHost h(c, V);
        if (!mbg::write_probtraj_csv(h.view(), V.m->cm, std::string(prefix) + "_probtraj.csv") ||
            !mbg::write_fp_csv(h.view(), V.m->cm, std::string(prefix) + "_fp.csv"))
            throw Error(MBG_ERR_ARG, std::string("cannot write output files with prefix ") + prefix);
    });
}

mbg_status mbg_write_variant_text(const mbg_ctx* c, int v, const char* path) {
    if (!c || !path) return MBG_ERR_ARG;
    return guarded(const_cast<mbg_ctx*>(c), [&] {
        const VariantRun& V = sweep_variant(c, v);
        const VariantHost h(c, V);
        if (std::string(path) == "-") {
            mbg::write_text(h.view(), V.m->cm, std::cout);
        } else {
            std::ofstream f(path);
            if (!f) throw Error(MBG_ERR_ARG, std::string("cannot open ") + path);
            mbg::write_text(h.view(), V.m->cm, f);
        }
    });
}

mbg_status mbg_get_variant_fixed_points(const mbg_ctx* c, int v, uint64_t* states, uint64_t* counts, size_t cap,
                                        size_t* n) {
    if (!c) return MBG_ERR_ARG;
    return guarded(const_cast<mbg_ctx*>(c), [&] {
        const VariantRun& V = sweep_variant(c, v);
        const size_t F = V.fixed_counts.size(), NW = (size_t)V.info.state_words;
        if (n) *n = F;
        const size_t k = std::min(cap, F);
        if (states && k) std::memcpy(states, V.fixed_states.data(), k * NW * 8);
        if (counts && k) std::memcpy(counts, V.fixed_counts.data(), k * 8);
    });
}

}  // extern "C"
