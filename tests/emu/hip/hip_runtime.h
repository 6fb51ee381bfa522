// CPU stand-in for <hip/hip_runtime.h> (test infrastructure): lets g++ compile the device code of
// 3d_navier_stokes_amd/csrc/nsg_kernels.hpp and nsg_fft.hpp unchanged, and run one thread block at
// a time with one host thread per GPU lane.  Lockstep is emulated at the synchronisation points the
// kernels state: a wavefront-scope barrier (__builtin_amdgcn_wave_barrier) is a barrier over the
// 64 lanes of that wavefront, __syncthreads over the block, and a cross-lane shuffle is a write to
// a per-wavefront buffer between two wavefront barriers.  A lane that reads LDS written by another
// lane without one of these in between reads whatever order the host threads produce -- which is
// exactly the class of missing-synchronisation mistake the emulation is there to expose (the
// kernels then disagree with the reference transform).  __shared__ arrays become function-static,
// i.e. shared by the threads of the one block that runs.
#pragma once
#include <algorithm>
#include <atomic>
#include <barrier>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>

struct double2 { double x, y; };
struct float2 { float x, y; };
inline double2 make_double2(double a, double b) { return double2{a, b}; }
inline float2 make_float2(float a, float b) { return float2{a, b}; }

namespace emu {
struct uint3v { unsigned x = 0, y = 0, z = 0; };
struct WaveCtx {
    std::barrier<>* bar = nullptr;
    uint64_t buf[64];
};
inline thread_local uint3v t_tid, t_bid, t_gdim, t_bdim;
inline thread_local WaveCtx* t_wave = nullptr;
inline thread_local std::barrier<>* t_block = nullptr;
inline std::mutex g_atomic;

inline void wave_barrier() { t_wave->bar->arrive_and_wait(); }
template <typename V> inline V shfl_from(V v, int src) {
    static_assert(sizeof(V) <= 8, "shuffle of at most 64 bits");
    const int lane = (int)(t_tid.x & 63);
    uint64_t w = 0;
    std::memcpy(&w, &v, sizeof(V));
    t_wave->buf[lane] = w;
    wave_barrier();
    const uint64_t r = t_wave->buf[src & 63];
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
inline void __syncthreads() { emu::t_block->arrive_and_wait(); }
template <typename V> inline V __shfl(V v, int src) { return emu::shfl_from(v, src); }
template <typename V> inline V __shfl_xor(V v, int mask) { return emu::shfl_from(v, (int)(emu::t_tid.x & 63) ^ mask); }
inline unsigned long long atomicMax(unsigned long long* p, unsigned long long v) {
    std::lock_guard<std::mutex> g(emu::g_atomic);
    const unsigned long long o = *p;
    *p = std::max(o, v);
    return o;
}
inline long long __double_as_longlong(double d) { long long r; std::memcpy(&r, &d, 8); return r; }
inline double __longlong_as_double(long long l) { double r; std::memcpy(&r, &l, 8); return r; }
