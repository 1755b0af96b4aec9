// libmbg.so: C ABI (include/mbg.h) over the front end, the model compiler, hipRTC
// code generation and the persistent trajectory kernel.
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <map>
#include <memory>
#include <numeric>
#include <sstream>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/mbg.h"
#include "front.h"
#include "model.h"
#include "output.h"

extern const char* const mbg_embedded_math;
extern const char* const mbg_embedded_kernel;

struct mbg_model {
    mbg::ParsedModel parsed;
    mbg::CompiledModel cm;
    std::vector<float> mult_up, mult_down, istate;  // empty: model's own values
};

namespace {

struct Error : std::runtime_error {
    mbg_status code;
    Error(mbg_status c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw Error(MBG_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

void copy_err(char* err, size_t n, const std::string& s) {
    if (err && n) {
        std::strncpy(err, s.c_str(), n - 1);
        err[n - 1] = 0;
    }
}

// Kernel argument block; layout identical to struct mbg_desc in device/mbg_kernel.hip.
struct Desc {
    const void* nodes;
    const void* inputs;
    const void* deps;
    const void* lut;
    const void* rates;
    const void* rtab;
    const void* gen_mask;
    const void* istate;
    const void* readouts;
    const void* ro_lg;
    void* prob;
    void* th;
    void* fin;
    void* fp_states;
    void* fp_count;
    void* summ;
    void* work_counter;
    void* flip_total;
    uint64_t fp_cap, summ_cap, nsteps, n_traj, first_id, seed;
    double maxt, tick, inv_tick, rscale;
    uint32_t variant_id;
    int N, K, W, FP, FT;
};

struct Shape {
    int nw, block, discrete, trace;
    bool operator<(const Shape& o) const {
        return std::tie(nw, block, discrete, trace) < std::tie(o.nw, o.block, o.discrete, o.trace);
    }
};

std::string generate_source(const mbg::CompiledModel& cm, const Shape& sh) {
    std::ostringstream os;
    os << "// Trajectory kernel generated for a model of " << cm.N << " nodes (" << cm.K
       << " non-internal), " << (cm.discrete ? "discrete" : "continuous") << " time.\n";
    os << "#define MBG_NW " << sh.nw << "\n";
    os << "#define MBG_BLOCK " << sh.block << "\n";
    os << "#define MBG_DISCRETE " << sh.discrete << "\n";
    os << "#define MBG_TRACE " << sh.trace << "\n";
    os << "#define MBG_HD __device__ __forceinline__\n";
    os << mbg_embedded_math << "\n" << mbg_embedded_kernel << "\n";
    return os.str();
}

struct Module {
    hipModule_t mod = nullptr;
    hipFunction_t fn = nullptr;
    int blocks_per_cu = 1;
};

template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    void alloc(size_t count) {
        n = count;
        hip_check(hipMalloc((void**)&p, std::max<size_t>(1, count) * sizeof(T)), "hipMalloc");
    }
    void upload(const std::vector<T>& v, hipStream_t s) {
        alloc(v.size());
        if (!v.empty()) hip_check(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s), "hipMemcpyAsync");
    }
    void zero(size_t count, hipStream_t s) {
        alloc(count);
        hip_check(hipMemsetAsync(p, 0, std::max<size_t>(1, count) * sizeof(T), s), "hipMemsetAsync");
    }
};

}  // namespace

struct mbg_ctx {
    int device = 0;
    int cus = 0;
    hipStream_t stream = nullptr;
    std::string err;
    std::map<Shape, Module> modules;
    // last run
    bool have = false;
    mbg_run_info info{};
    std::vector<uint64_t> prob, th, fin, fixed_states, fixed_counts, summ;
};

namespace {

const Module& get_module(mbg_ctx* c, const mbg::CompiledModel& cm, const Shape& sh) {
    auto it = c->modules.find(sh);
    if (it != c->modules.end()) return it->second;
    const std::string src = generate_source(cm, sh);
    hiprtcProgram prog;
    if (hiprtcCreateProgram(&prog, src.c_str(), "mbg_kernel.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS)
        throw Error(MBG_ERR_COMPILE, "hiprtcCreateProgram failed");
    const char* opts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off",
                          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
    const hiprtcResult rc = hiprtcCompileProgram(prog, (int)(sizeof(opts) / sizeof(opts[0])), opts);
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
    hip_check(hipModuleGetFunction(&m.fn, m.mod, "mbg_simulate"), "hipModuleGetFunction");
    int nb = 0;
    if (hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&nb, m.fn, sh.block, 0) != hipSuccess || nb < 1) nb = 1;
    m.blocks_per_cu = nb;
    return c->modules.emplace(sh, m).first->second;
}

int default_block(const mbg::CompiledModel&) { return 64; }

void run_impl(mbg_ctx* c, const mbg_model* m, const mbg_run_params* p) {
    const mbg::CompiledModel& cm = m->cm;
    if (cm.K > mbg::kMaxReadouts) throw Error(MBG_ERR_LIMIT, "This executable supports a maximum of 20 non-internal nodes.");
    if (!(cm.time_tick > 0.f) || !(cm.max_time > 0.f)) throw Error(MBG_ERR_ARG, "max_time and time_tick must be positive");
    const uint64_t n = p->n_trajectories ? p->n_trajectories : (uint64_t)cm.sample_count;
    const uint64_t count = p->count ? p->count : n;
    if (n == 0 || count == 0) throw Error(MBG_ERR_ARG, "no trajectories to simulate");
    if (p->first_id + count > (1ull << 32)) throw Error(MBG_ERR_ARG, "trajectory ids must stay below 2^32");
    const int block = p->block ? p->block : default_block(cm);
    if (block <= 0 || block % 64 || block > 1024) throw Error(MBG_ERR_ARG, "block must be a positive multiple of 64, at most 1024");
    const Shape sh{cm.NW, block, cm.discrete ? 1 : 0, p->trace ? 1 : 0};
    const Module& mod = get_module(c, cm, sh);

    const mbg::RateTables rt = mbg::make_rate_tables(cm, m->mult_up.empty() ? nullptr : m->mult_up.data(),
                                                     m->mult_down.empty() ? nullptr : m->mult_down.data(),
                                                     m->istate.empty() ? nullptr : m->istate.data());
    const int W = (int)std::ceil(cm.max_time / cm.time_tick);  // window_average_small.cpp:26
    const uint64_t S = 1ull << cm.K;
    int FP, FT;
    if (cm.discrete) {
        FP = 0;
        FT = 59 - mbg::ceil_log2_u64(n);
    } else {
        const int eq = mbg::ceil_log2_real((double)n * (double)cm.time_tick);
        FP = 60 - eq;
        FT = 57 - eq;
    }
    const float ratio = cm.max_time / cm.time_tick;
    const uint64_t nsteps = ratio >= 2.f ? (uint64_t)std::floor(ratio) : 1;
    // Fixed points: one slot per simulated trajectory, within a 16 GiB budget.
    const uint64_t fp_cap = std::min<uint64_t>(count, (16ull << 30) / (8ull * (uint64_t)cm.NW));

    hipStream_t s = c->stream;
    DevBuf<mbg::NodeRec> d_nodes;
    DevBuf<uint32_t> d_inputs, d_deps, d_lut, d_ro;
    DevBuf<uint64_t> d_rates, d_rtab, d_gm, d_prob, d_th, d_fin, d_fp, d_fpc, d_summ, d_work, d_flips;
    DevBuf<float> d_istate;
    DevBuf<double> d_rolg;
    d_nodes.upload(cm.nodes, s);
    d_inputs.upload(cm.inputs, s);
    d_deps.upload(cm.deps, s);
    d_lut.upload(cm.lut, s);
    std::vector<uint32_t> ro(cm.readouts.begin(), cm.readouts.end());
    d_ro.upload(ro, s);
    d_rates.upload(rt.rates, s);
    d_rtab.upload(rt.rtab, s);
    d_gm.upload(cm.gen_mask, s);
    d_istate.upload(rt.istate, s);
    d_rolg.upload(rt.ro_lg, s);
    d_prob.zero((size_t)W * S, s);
    d_th.zero((size_t)W, s);
    d_fin.zero((size_t)S, s);
    d_fp.alloc((size_t)fp_cap * cm.NW);
    d_fpc.zero(1, s);
    const uint64_t summ_cap = p->trace ? count : 0;
    d_summ.zero((size_t)summ_cap * 4, s);
    d_work.zero(1, s);
    d_flips.zero(1, s);

    Desc D{};
    D.nodes = d_nodes.p;
    D.inputs = d_inputs.p;
    D.deps = d_deps.p;
    D.lut = d_lut.p;
    D.rates = d_rates.p;
    D.rtab = d_rtab.p;
    D.gen_mask = d_gm.p;
    D.istate = d_istate.p;
    D.readouts = d_ro.p;
    D.ro_lg = d_rolg.p;
    D.prob = d_prob.p;
    D.th = d_th.p;
    D.fin = d_fin.p;
    D.fp_states = d_fp.p;
    D.fp_count = d_fpc.p;
    D.summ = d_summ.p;
    D.work_counter = d_work.p;
    D.flip_total = d_flips.p;
    D.fp_cap = fp_cap;
    D.summ_cap = summ_cap;
    D.nsteps = nsteps;
    D.n_traj = count;
    D.first_id = p->first_id;
    D.seed = p->use_model_seed ? (uint64_t)cm.seed : p->seed;
    D.maxt = (double)cm.max_time;
    D.tick = (double)cm.time_tick;
    D.inv_tick = 1.0 / D.tick;
    D.rscale = std::ldexp(1.0, -rt.E);
    D.variant_id = p->variant_id;
    D.N = cm.N;
    D.K = cm.K;
    D.W = W;
    D.FP = FP;
    D.FT = FT;

    uint64_t grid = p->grid ? (uint64_t)p->grid : (uint64_t)c->cus * (uint64_t)mod.blocks_per_cu;
    grid = std::max<uint64_t>(1, std::min<uint64_t>(grid, (count + block - 1) / block));
    void* args[] = {&D};
    hipEvent_t e0, e1;
    hip_check(hipEventCreate(&e0), "hipEventCreate");
    hip_check(hipEventCreate(&e1), "hipEventCreate");
    hip_check(hipEventRecord(e0, s), "hipEventRecord");
    hip_check(hipModuleLaunchKernel(mod.fn, (unsigned)grid, 1, 1, (unsigned)block, 1, 1, 0, s, args, nullptr),
              "hipModuleLaunchKernel");
    hip_check(hipEventRecord(e1, s), "hipEventRecord");
    hip_check(hipStreamSynchronize(s), "kernel execution");
    float ms = 0.f;
    hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);

    c->prob.resize((size_t)W * S);
    c->th.resize(W);
    c->fin.resize(S);
    hip_check(hipMemcpy(c->prob.data(), d_prob.p, c->prob.size() * 8, hipMemcpyDeviceToHost), "hipMemcpy");
    hip_check(hipMemcpy(c->th.data(), d_th.p, c->th.size() * 8, hipMemcpyDeviceToHost), "hipMemcpy");
    hip_check(hipMemcpy(c->fin.data(), d_fin.p, c->fin.size() * 8, hipMemcpyDeviceToHost), "hipMemcpy");
    uint64_t nfp = 0, flips = 0;
    hip_check(hipMemcpy(&nfp, d_fpc.p, 8, hipMemcpyDeviceToHost), "hipMemcpy");
    hip_check(hipMemcpy(&flips, d_flips.p, 8, hipMemcpyDeviceToHost), "hipMemcpy");
    if (nfp > fp_cap) throw Error(MBG_ERR_LIMIT, "fixed-point buffer overflow");
    std::vector<uint64_t> fps((size_t)nfp * cm.NW);
    if (nfp) hip_check(hipMemcpy(fps.data(), d_fp.p, fps.size() * 8, hipMemcpyDeviceToHost), "hipMemcpy");
    c->summ.resize((size_t)summ_cap * 4);
    if (summ_cap) hip_check(hipMemcpy(c->summ.data(), d_summ.p, c->summ.size() * 8, hipMemcpyDeviceToHost), "hipMemcpy");

    // distinct fixed points in the reference order (src/state.h:43-49)
    const int NW = cm.NW;
    std::vector<uint64_t> idx(nfp);
    std::iota(idx.begin(), idx.end(), 0);
    auto row = [&](uint64_t r) { return fps.data() + r * NW; };
    std::sort(idx.begin(), idx.end(), [&](uint64_t a, uint64_t b) {
        const uint64_t *x = row(a), *y = row(b);
        for (int w = NW - 1; w >= 0; w--)
            if (x[w] != y[w]) return x[w] < y[w];
        return false;
    });
    c->fixed_states.clear();
    c->fixed_counts.clear();
    for (size_t i = 0; i < idx.size();) {
        size_t j = i;
        while (j < idx.size() && std::equal(row(idx[i]), row(idx[i]) + NW, row(idx[j]))) j++;
        c->fixed_states.insert(c->fixed_states.end(), row(idx[i]), row(idx[i]) + NW);
        c->fixed_counts.push_back(j - i);
        i = j;
    }

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
    I.kernel_ms = ms;
    I.time_tick = cm.time_tick;
    I.max_time = cm.max_time;
    c->have = true;
}

mbg::Results results_view(const mbg_ctx* c) {
    mbg::Results r;
    r.info = &c->info;
    r.prob = &c->prob;
    r.th = &c->th;
    r.fin = &c->fin;
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
        mbg::parse_model(bnd, cfg, m->parsed);
        m->cm = mbg::compile_model(m->parsed);
    } catch (const mbg::ParseError& e) {
        copy_err(err, errlen, e.what());
        return MBG_ERR_PARSE;
    } catch (const std::exception& e) {
        copy_err(err, errlen, e.what());
        return MBG_ERR_PARSE;
    }
    if (m->cm.K > mbg::kMaxReadouts) {
        copy_err(err, errlen, "This executable supports a maximum of 20 non-internal nodes.");
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
    auto it = m->parsed.constants.find(name);
    if (it == m->parsed.constants.end()) return MBG_ERR_ARG;
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

mbg_status mbg_model_generate_source(const mbg_model* m, int trace, char* buf, size_t cap, size_t* len) {
    if (!m) return MBG_ERR_ARG;
    const std::string s = generate_source(m->cm, Shape{m->cm.NW, default_block(m->cm), m->cm.discrete ? 1 : 0, trace ? 1 : 0});
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
    if (hipStreamCreate(&c->stream) != hipSuccess) return MBG_ERR_HIP;
    *out = c.release();
    return MBG_OK;
}

void mbg_destroy(mbg_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    for (auto& kv : c->modules) (void)hipModuleUnload(kv.second.mod);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char* mbg_last_error(const mbg_ctx* c) { return c ? c->err.c_str() : "null context"; }

mbg_status mbg_compile(mbg_ctx* c, const mbg_model* m, int trace) {
    if (!c || !m) return MBG_ERR_ARG;
    return guarded(c, [&] {
        hip_check(hipSetDevice(c->device), "hipSetDevice");
        get_module(c, m->cm, Shape{m->cm.NW, default_block(m->cm), m->cm.discrete ? 1 : 0, trace ? 1 : 0});
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

mbg_status mbg_get_run_info(const mbg_ctx* c, mbg_run_info* out) {
    if (!c || !out) return MBG_ERR_ARG;
    if (!c->have) return MBG_ERR_NOT_BUILT;
    *out = c->info;
    return MBG_OK;
}

mbg_status mbg_get_probtraj_raw(const mbg_ctx* c, uint64_t* prob, uint64_t* th) {
    if (!c) return MBG_ERR_ARG;
    if (!c->have) return MBG_ERR_NOT_BUILT;
    if (prob) std::memcpy(prob, c->prob.data(), c->prob.size() * 8);
    if (th) std::memcpy(th, c->th.data(), c->th.size() * 8);
    return MBG_OK;
}

mbg_status mbg_get_probtraj(const mbg_ctx* c, double* prob, double* th) {
    if (!c) return MBG_ERR_ARG;
    if (!c->have) return MBG_ERR_NOT_BUILT;
    const mbg::Results r = results_view(c);
    if (prob)
        for (size_t i = 0; i < c->prob.size(); i++) prob[i] = mbg::window_prob(r, i);
    if (th)
        for (size_t w = 0; w < c->th.size(); w++) th[w] = mbg::window_th(r, w);
    return MBG_OK;
}

mbg_status mbg_get_final_states(const mbg_ctx* c, uint64_t* counts) {
    if (!c || !counts) return MBG_ERR_ARG;
    if (!c->have) return MBG_ERR_NOT_BUILT;
    std::memcpy(counts, c->fin.data(), c->fin.size() * 8);
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
