// CPU stand-in for <hip/hip_runtime.h> (test infrastructure): lets g++ compile the device code of
// 3d_navier_stokes_amd/csrc/nsg_kernels.hpp and nsg_fft.hpp unchanged, and run it on the host.
//
// Execution model: the kernels it serves (the persistent z passes) synchronise only inside a
// wavefront, so a block's wavefronts run one after another, and the 64 lanes of a wavefront run
// as cooperative coroutines (ucontext) on one host thread.  A lane runs until it reaches a
// wavefront synchronisation point -- __builtin_amdgcn_wave_barrier (the kernels' LDS exchange
// barrier) or a cross-lane shuffle -- and then hands over to the next lane; when the last lane
// arrives, the first resumes.  So every lane has executed everything before barrier k before any
// lane executes anything after it, which is the guarantee the kernels rely on, and nothing more:
// a lane that reads LDS written by another lane with no barrier in between sees whatever the
// round-robin order produced, which is the kind of missing-synchronisation mistake this is meant
// to expose (the result then disagrees with the reference transform).  __shared__ arrays become
// function-static, i.e. shared by the lanes of the one wavefront that runs.  __syncthreads is not
// supported (it would need every wavefront of the block to interleave) and aborts.
#pragma once
#include <ucontext.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <vector>

struct double2 { double x, y; };
struct float2 { float x, y; };
inline double2 make_double2(double a, double b) { return double2{a, b}; }
inline float2 make_float2(float a, float b) { return float2{a, b}; }

namespace emu {
struct uint3v { unsigned x = 0, y = 0, z = 0; };
inline uint3v t_tid, t_bid, t_gdim, t_bdim;

struct Wave {
    static constexpr size_t kStack = 1 << 18;
    ucontext_t main;
    ucontext_t ctx[64];
    std::vector<char> stacks = std::vector<char>(64 * kStack);
    bool done[64];
    int lanes = 64, cur = 0, first_tid = 0;
    uint64_t buf[64];
    std::function<void()> body;
};
inline Wave* g_wave = nullptr;

inline void enter(int lane) {
    g_wave->cur = lane;
    t_tid.x = (unsigned)(g_wave->first_tid + lane);
}
// hand over to the next lane that has not finished (round-robin); back to the host when all have
inline void next_from(int me) {
    Wave& w = *g_wave;
    for (int k = 1; k <= w.lanes; ++k) {
        const int nx = (me + k) % w.lanes;
        if (!w.done[nx]) {
            if (nx == me) return;
            enter(nx);
            swapcontext(&w.ctx[me], &w.ctx[nx]);
            return;
        }
    }
    swapcontext(&w.ctx[me], &w.main);
}
inline void lane_entry() {
    Wave& w = *g_wave;
    w.body();
    const int me = w.cur;
    w.done[me] = true;
    for (int k = 1; k < w.lanes; ++k) {
        const int nx = (me + k) % w.lanes;
        if (!w.done[nx]) {
            enter(nx);
            setcontext(&w.ctx[nx]);
        }
    }
    setcontext(&w.main);
}
// run body() as `lanes` lanes of one wavefront whose first thread index is first_tid
inline void run_wave(Wave& w, int first_tid, int lanes, std::function<void()> body) {
    g_wave = &w;
    w.lanes = lanes;
    w.first_tid = first_tid;
    w.body = std::move(body);
    for (int l = 0; l < lanes; ++l) {
        w.done[l] = false;
        getcontext(&w.ctx[l]);
        w.ctx[l].uc_stack.ss_sp = w.stacks.data() + (size_t)l * Wave::kStack;
        w.ctx[l].uc_stack.ss_size = Wave::kStack;
        w.ctx[l].uc_link = nullptr;
        makecontext(&w.ctx[l], (void (*)())lane_entry, 0);
    }
    enter(0);
    swapcontext(&w.main, &w.ctx[0]);
}

inline void wave_barrier() { next_from(g_wave->cur); }
template <typename V> inline V shfl_from(V v, int src) {
    static_assert(sizeof(V) <= 8, "shuffle of at most 64 bits");
    uint64_t x = 0;
    std::memcpy(&x, &v, sizeof(V));
    g_wave->buf[g_wave->cur] = x;
    wave_barrier();
    const uint64_t r = g_wave->buf[src & 63];
    wave_barrier();
    V out;
    std::memcpy(&out, &r, sizeof(V));
    return out;
}
}  // namespace emu

#define threadIdx emu::t_tid
#define blockIdx emu::t_bid
#define gridDim emu::t_gdim
#define blockDim emu::t_bdim
#define __device__
#define __host__
#define __global__
#define __forceinline__ inline
#define __shared__ static
#define __launch_bounds__(...)
#define __builtin_amdgcn_fence(order, scope) ((void)0)
#define __builtin_amdgcn_wave_barrier() emu::wave_barrier()
inline void __syncthreads() {
    std::fprintf(stderr, "emulation: __syncthreads is not supported\n");
    std::abort();
}
template <typename V> inline V __shfl(V v, int src) { return emu::shfl_from(v, src); }
// global->LDS copies (pass 5's pipelined kernel) are not emulated: the z passes do not use them
inline void __builtin_amdgcn_global_load_lds(const void*, void*, unsigned, int, int) {
    std::fprintf(stderr, "emulation: global->LDS copies are not supported\n");
    std::abort();
}
// the value of the first lane (every lane of an emulated wavefront is active)
inline int __builtin_amdgcn_readfirstlane(int v) { return emu::shfl_from(v, 0); }
template <typename V> inline V __shfl_xor(V v, int mask) { return emu::shfl_from(v, (int)(emu::t_tid.x & 63) ^ mask); }
inline unsigned long long atomicMax(unsigned long long* p, unsigned long long v) {
    const unsigned long long o = *p;
    *p = std::max(o, v);
    return o;
}
inline long long __double_as_longlong(double d) { long long r; std::memcpy(&r, &d, 8); return r; }
inline double __longlong_as_double(long long l) { double r; std::memcpy(&r, &l, 8); return r; }
