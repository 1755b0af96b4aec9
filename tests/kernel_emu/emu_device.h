// Host emulation of the HIP device environment the trajectory kernel is written against, so
// that the kernel source itself (upmabossg_amd/csrc/device/mbg_kernel.hip) can be compiled with
// g++ and run on the CPU as test infrastructure.
//
// Every GPU thread of a launch becomes a coroutine (ucontext) with its own stack.  One host
// thread runs them round-robin; a coroutine runs until it reaches a collective operation and
// then yields:
//   - wave collectives (__ballot, __shfl, __builtin_amdgcn_wave_barrier) complete once every
//     lane of the 64-lane wave that has not returned has reached one; __ballot's mask and
//     __shfl's source read only those lanes;
//   - __syncthreads completes once every thread of the block that has not returned is waiting.
// Between collectives the lanes of a wave run one after another, which is one legal schedule of
// a wave whose lanes diverge.  Atomics are plain read-modify-writes (one host thread).
// __builtin_amdgcn_readfirstlane returns its argument: the kernel applies it to wave-uniform
// values only.
//
// Kernels with static LDS (__shared__ locals) run one block at a time, since the locals are
// shared by every block (static storage); kernels that use only the dynamic LDS may run their
// blocks interleaved, each with its own exactly sized LDS allocation.  Under AddressSanitizer
// (g++ -fsanitize=address, host code) every emulated buffer is bounds-checked; the coroutine
// switches are announced to it with the sanitizer's fiber interface.
#pragma once

#include <ucontext.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

#if defined(__SANITIZE_ADDRESS__)
#include <sanitizer/common_interface_defs.h>
#define EMU_ASAN 1
#else
#define EMU_ASAN 0
#endif

// Under MemorySanitizer: memory the kernel must write before reading it (LDS, device
// buffers) is marked never-written, and inputs read from files (written by uninstrumented
// library code) are marked written.
#if defined(__has_feature)
#if __has_feature(memory_sanitizer)
#include <sanitizer/msan_interface.h>
#define EMU_MSAN 1
#endif
#endif
#ifdef EMU_MSAN
inline void emu_poison(const void* p, size_t n) { __msan_poison(p, n); }
inline void emu_unpoison(const void* p, size_t n) { __msan_unpoison(p, n); }
#else
inline void emu_poison(const void*, size_t) {}
inline void emu_unpoison(const void*, size_t) {}
#endif

namespace emu {

struct Dim3 {
    unsigned x = 0, y = 0, z = 0;
};

enum class St { Run, Wave, Bar, Done };
enum Op { OpBallot = 1, OpShfl = 2, OpWaveBar = 3 };

struct Block {
    Dim3 idx;
    std::vector<unsigned long long> lds;  // dynamic LDS of this block
};

struct Thread {
    ucontext_t ctx;
    char* stack = nullptr;
    size_t stack_size = 0;
    Dim3 tid;
    Block* blk = nullptr;
    St st = St::Run;
    int op = 0, src = 0;
    uint64_t arg = 0, res = 0;
    void* fake_stack = nullptr;  // AddressSanitizer fiber bookkeeping
};

struct World {
    ucontext_t sched;
    Thread* cur = nullptr;
    Dim3 grid, block;
    std::function<void()> body;
    const void* sched_bottom = nullptr;
    size_t sched_size = 0;
    void* sched_fake = nullptr;
    std::string error;
};

inline World& W() {
    static World w;
    return w;
}

// Switches from the running coroutine back to the scheduler.
inline void yield_to_sched() {
    World& w = W();
    Thread* t = w.cur;
#if EMU_ASAN
    __sanitizer_start_switch_fiber(&t->fake_stack, w.sched_bottom, w.sched_size);
#endif
    swapcontext(&t->ctx, &w.sched);
#if EMU_ASAN
    __sanitizer_finish_switch_fiber(t->fake_stack, &w.sched_bottom, &w.sched_size);
#endif
}

inline void trampoline() {
    World& w = W();
#if EMU_ASAN
    __sanitizer_finish_switch_fiber(nullptr, &w.sched_bottom, &w.sched_size);
#endif
    w.body();
    w.cur->st = St::Done;
#if EMU_ASAN
    __sanitizer_start_switch_fiber(nullptr, w.sched_bottom, w.sched_size);  // this fiber ends
#endif
    // returning switches to uc_link (the scheduler)
}

inline uint64_t wave_op(int op, uint64_t arg, int src) {
    Thread* t = W().cur;
    t->st = St::Wave;
    t->op = op;
    t->arg = arg;
    t->src = src;
    yield_to_sched();
    return t->res;
}

inline void block_barrier() {
    W().cur->st = St::Bar;
    yield_to_sched();
}

// Runs one launch: grid x block threads executing `body`.  interleave: all blocks' threads
// are scheduled together (kernels without static LDS); otherwise one block after another.
// smem_bytes: dynamic LDS per block.  Throws on a deadlock (collectives that can never
// complete, such as a wave whose lanes wait at different operations).
inline void launch(Dim3 grid, Dim3 block, size_t smem_bytes, bool interleave, std::function<void()> body) {
    World& w = W();
    w.grid = grid;
    w.block = block;
    w.body = std::move(body);
    const unsigned nthreads = block.x * block.y * block.z;
    if (block.y != 1 || block.z != 1 || nthreads % 64 != 0)
        throw std::runtime_error("emu: blocks must be one-dimensional multiples of 64 threads");
    const size_t stack_size = EMU_ASAN ? (512u << 10) : (128u << 10);
    std::vector<std::pair<unsigned, unsigned>> all;  // (bx, by)
    for (unsigned by = 0; by < grid.y; by++)
        for (unsigned bx = 0; bx < grid.x; bx++) all.push_back({bx, by});
    const size_t group = interleave ? all.size() : 1;
    for (size_t g0 = 0; g0 < all.size(); g0 += group) {
        const size_t g1 = std::min(all.size(), g0 + group);
        std::vector<Block> blocks(g1 - g0);
        std::vector<Thread> threads(blocks.size() * nthreads);
        for (size_t b = 0; b < blocks.size(); b++) {
            blocks[b].idx = Dim3{all[g0 + b].first, all[g0 + b].second, 0};
            blocks[b].lds.assign((smem_bytes + 7) / 8, 0xA5A5A5A5A5A5A5A5ull);  // LDS is not zeroed
            emu_poison(blocks[b].lds.data(), blocks[b].lds.size() * 8);
            for (unsigned i = 0; i < nthreads; i++) {
                Thread& t = threads[b * nthreads + i];
                t.tid = Dim3{i, 0, 0};
                t.blk = &blocks[b];
                t.stack = static_cast<char*>(std::malloc(stack_size));
                t.stack_size = stack_size;
                getcontext(&t.ctx);
                t.ctx.uc_stack.ss_sp = t.stack;
                t.ctx.uc_stack.ss_size = stack_size;
                t.ctx.uc_link = &w.sched;
                makecontext(&t.ctx, (void (*)())trampoline, 0);
            }
        }
        const size_t nwaves = threads.size() / 64;
        for (;;) {
            bool progressed = false;
            for (Thread& t : threads) {
                if (t.st != St::Run) continue;
                w.cur = &t;
#if EMU_ASAN
                __sanitizer_start_switch_fiber(&w.sched_fake, t.stack, t.stack_size);
#endif
                swapcontext(&w.sched, &t.ctx);
#if EMU_ASAN
                __sanitizer_finish_switch_fiber(w.sched_fake, nullptr, nullptr);
#endif
                progressed = true;
            }
            bool resolved = false, live = false;
            for (size_t wv = 0; wv < nwaves; wv++) {
                Thread* lanes = &threads[wv * 64];
                int waiting = 0, alive = 0, op = 0;
                bool mixed = false;
                for (int l = 0; l < 64; l++) {
                    if (lanes[l].st == St::Done) continue;
                    alive++;
                    if (lanes[l].st == St::Wave) {
                        if (op && lanes[l].op != op) mixed = true;
                        op = lanes[l].op;
                        waiting++;
                    }
                }
                if (alive) live = true;
                if (!alive || waiting != alive) continue;
                if (mixed) throw std::runtime_error("emu: lanes of one wave wait at different collectives");
                uint64_t mask = 0;
                for (int l = 0; l < 64; l++)
                    if (lanes[l].st == St::Wave && op == OpBallot && lanes[l].arg) mask |= 1ull << l;
                for (int l = 0; l < 64; l++) {  // results first, from the waiting lanes' arguments
                    Thread& t = lanes[l];
                    if (t.st != St::Wave) continue;
                    if (op == OpBallot) {
                        t.res = mask;
                    } else if (op == OpShfl) {
                        const int s = t.src & 63;
                        if (lanes[s].st != St::Wave)
                            throw std::runtime_error("emu: __shfl reads a lane that has returned");
                        t.res = lanes[s].arg;
                    } else {
                        t.res = 0;
                    }
                }
                for (int l = 0; l < 64; l++)
                    if (lanes[l].st == St::Wave) lanes[l].st = St::Run;
                resolved = true;
            }
            for (size_t b = 0; b < blocks.size(); b++) {
                Thread* bt = &threads[b * nthreads];
                unsigned alive = 0, bar = 0;
                for (unsigned i = 0; i < nthreads; i++) {
                    if (bt[i].st == St::Done) continue;
                    alive++;
                    if (bt[i].st == St::Bar) bar++;
                }
                if (alive && bar == alive) {
                    for (unsigned i = 0; i < nthreads; i++)
                        if (bt[i].st == St::Bar) bt[i].st = St::Run;
                    resolved = true;
                }
            }
            if (!live) break;
            if (!progressed && !resolved) throw std::runtime_error("emu: deadlock (collectives that cannot complete)");
        }
        for (Thread& t : threads) std::free(t.stack);
    }
    w.cur = nullptr;
}

inline unsigned long long* dyn_lds() { return W().cur->blk->lds.data(); }

inline uint32_t bitreverse32(uint32_t v) {
    v = ((v >> 1) & 0x55555555u) | ((v & 0x55555555u) << 1);
    v = ((v >> 2) & 0x33333333u) | ((v & 0x33333333u) << 2);
    v = ((v >> 4) & 0x0F0F0F0Fu) | ((v & 0x0F0F0F0Fu) << 4);
    v = ((v >> 8) & 0x00FF00FFu) | ((v & 0x00FF00FFu) << 8);
    return (v >> 16) | (v << 16);
}

template <class T>
inline T shfl(T v, int src) {
    static_assert(sizeof(T) <= 8, "emu: __shfl of a value wider than 64 bits");
    uint64_t a = 0;
    std::memcpy(&a, &v, sizeof(T));
    const uint64_t r = wave_op(OpShfl, a, src);
    T out;
    std::memcpy(&out, &r, sizeof(T));
    return out;
}

}  // namespace emu

// ---------------------------------------------------------------- the HIP names the kernel uses
#define __device__
#define __forceinline__ inline
#define __global__
#define __launch_bounds__(...)
#define __shared__ static
#define CONST_AS
#define MBG_OPAQUE_V(x) ((void)0)
#define MBG_OPAQUE_S(x) ((void)0)
#define MBG_DYN_LDS(name) u64* name = emu::dyn_lds()
// four u32 read as one 16-byte access (GNU vector; indexed like the device's ext_vector_type)
#define MBG_U32X4 1
typedef unsigned int mbg_u32x4 __attribute__((vector_size(16)));

#define threadIdx (emu::W().cur->tid)
#define blockIdx (emu::W().cur->blk->idx)
#define blockDim (emu::W().block)
#define gridDim (emu::W().grid)

#define __syncthreads() emu::block_barrier()
#define __ballot(p) emu::wave_op(emu::OpBallot, (p) ? 1ull : 0ull, 0)
#define __shfl(v, s) emu::shfl((v), (s))
#define __builtin_amdgcn_wave_barrier() ((void)emu::wave_op(emu::OpWaveBar, 0, 0))
#define __builtin_amdgcn_fence(...) ((void)0)
#define __builtin_amdgcn_readfirstlane(v) (v)
#define __builtin_bitreverse32(v) emu::bitreverse32(v)
#define __ffsll(v) __builtin_ffsll(v)
#define __popcll(v) __builtin_popcountll(v)
#define __HIP_MEMORY_SCOPE_AGENT 0
#define __hip_atomic_load(p, order, scope) (*(p))

template <class T>
inline T atomicAdd(T* p, T v) {
    const T old = *p;
    *p = old + v;
    return old;
}
template <class T>
inline T atomicCAS(T* p, T cmp, T v) {
    const T old = *p;
    if (old == cmp) *p = v;
    return old;
}
template <class T>
inline T atomicMin(T* p, T v) {
    const T old = *p;
    if (v < old) *p = v;
    return old;
}
template <class T>
inline T atomicMax(T* p, T v) {
    const T old = *p;
    if (v > old) *p = v;
    return old;
}
// the device's constant-rate clock: not modelled (the kernels only time themselves with it)
inline unsigned long long wall_clock64() { return 0; }
using std::max;
using std::min;
