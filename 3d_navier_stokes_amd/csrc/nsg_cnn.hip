// nsg_cnn.hip -- BF16 MFMA kernels of the learned subgrid-stress network: convolution forward
// (implicit GEMM on v_mfma_f32_16x16x32_bf16), input gradients (the same kernel on flipped,
// transposed weights), weight gradients, MSE loss and Adam; the nsg_cnn_* C ABI.
//
// Convolution tile: 4 x 8 x 16 output points (x, y, z; z fastest) per 512-thread block, the
// 6 x 10 x 18 periodic halo of the input staged once in LDS (point stride CIN + 8 halfwords so
// that the 16 points of one MFMA row block fall in distinct banks).  GEMM view per tile:
// M = 512 points (32 row blocks of one z-row each, 4 per wavefront), N = output channels,
// K = 27 taps x CIN input channels, consumed 32 at a time.  MFMA operand layout (16x16x32 bf16):
// lane l holds A[l % 16][8 (l / 16) + j] and B[8 (l / 16) + j][l % 16], j = 0..7, and receives
// D[4 (l / 16) + i][l % 16], i = 0..3.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/nsg.h"
#include "nsg_cnn.hpp"

using namespace nsgcnn;

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float floatx4;
typedef __attribute__((__vector_size__(4 * sizeof(unsigned int)))) unsigned int u32x4;

namespace {

constexpr int TXT = 4, TYT = 8, TZT = 16;
constexpr int HX = TXT + 2, HY = TYT + 2, HZ = TZT + 2;
constexpr int HPTS = HX * HY * HZ;
constexpr int TPTS = TXT * TYT * TZT;
constexpr int CONV_THREADS = 512;

__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);  // NaN stays NaN
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
// the same rounding (to nearest, ties to even) by gfx950's conversion instruction
__device__ __forceinline__ uint16_t f2bf_hw(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }

// GELU of the forward convolution's epilogue, x Phi(x) = max(x, 0) - |x| erfc(|x| / sqrt 2) / 2,
// with erfc from Abramowitz & Stegun 7.1.26 (erfc(z) = t (a1 + t (a2 + t (a3 + t (a4 + t a5))))
// exp(-z^2), t = 1 / (1 + p z), absolute error <= 1.5e-7) written in w = x sqrt(log2(e) / 2), so
// that exp(-z^2) = exp2(-w^2) and p z = p' |w| with p' = p / sqrt(log2 e); the halves are folded
// into the coefficients.  One branch-free path of 14 instructions (hardware reciprocal and
// exponential); its difference from the erff form is far below half a bfloat16 unit of the
// rounded activation.
__device__ __forceinline__ float gelu_as(float x) {
    const float w = x * 0.84932180028801907f;
    const float t = __builtin_amdgcn_rcpf(fmaf(0.27273748087922250f, fabsf(w), 1.0f));
    const float q = t * fmaf(t, fmaf(t, fmaf(t, fmaf(t, 0.5307027145f, -0.7265760135f), 0.7107068705f), -0.142248368f),
                             0.127414796f);
    const float e = __builtin_amdgcn_exp2f(-(w * w));
    return fmaf(-fabsf(x), q * e, fmaxf(x, 0.0f));
}
__device__ __forceinline__ float gelu_grad(float x) {
    return 0.5f * (1.0f + erff(x * 0.70710678118654752f)) + x * 0.39894228040143268f * expf(-0.5f * x * x);
}

__device__ __forceinline__ bf16x8 ld16(const uint16_t* p) { return __builtin_bit_cast(bf16x8, *(const u32x4*)p); }

// 16 bytes per lane from global memory to LDS at lds + 16 lane (lds wave-uniform), written in
// assembly so that the compiler does not track it: its own wait insertion treats any such copy as
// possibly aliasing every LDS read and would make the MFMA loop wait for the copies of the next
// item; the kernel waits for them itself.  No compiler-generated code in the kernels that use it
// keeps a value in M0.
__device__ __forceinline__ void dma16(const void* g, const uint16_t* lds) {
    const uint32_t off = (uint32_t)(uintptr_t)(__attribute__((address_space(3))) const void*)lds;
    asm volatile("s_mov_b32 m0, %1\n\tglobal_load_lds_dwordx4 %0, off" : : "v"(g), "s"(off) : "memory");
}

// the same copy from base + off bytes (base wave-uniform, off < 2^32): the address is formed by
// the instruction (scalar base, 32-bit vector offset)
__device__ __forceinline__ void dma16o(const void* base, unsigned off, const uint16_t* lds) {
    const uint32_t l = (uint32_t)(uintptr_t)(__attribute__((address_space(3))) const void*)lds;
    asm volatile("s_mov_b32 m0, %2\n\tglobal_load_lds_dwordx4 %0, %1" : : "v"(off), "s"(base), "s"(l) : "memory");
}

__device__ __forceinline__ int wrap(int v, int N) { return v < 0 ? v + N : (v >= N ? v - N : v); }

// decompose a tile index into its origin; tiles ordered z fastest
__device__ __forceinline__ void tile_origin(int t, int N, int& x0, int& y0, int& z0) {
    const int ntz = N / TZT, nty = N / TYT;
    z0 = (t % ntz) * TZT;
    t /= ntz;
    y0 = (t % nty) * TYT;
    x0 = (t / nty) * TXT;
}

// stage the periodic halo of a tile (channels-last, CIN channels) into LDS with point stride SP
template <int CIN, int SP>
__device__ __forceinline__ void stage_halo(uint16_t* lds, const uint16_t* in, int N, int x0, int y0, int z0) {
    constexpr int CH = CIN / 8;
    for (int i = threadIdx.x; i < HPTS * CH; i += blockDim.x) {
        const int q = i / CH, c = i % CH;
        const int hz = q % HZ, hy = (q / HZ) % HY, hx = q / (HZ * HY);
        const int gx = wrap(x0 + hx - 1, N), gy = wrap(y0 + hy - 1, N), gz = wrap(z0 + hz - 1, N);
        const long long gp = ((long long)gx * N + gy) * N + gz;
        *(u32x4*)(lds + q * SP + c * 8) = *(const u32x4*)(in + gp * CIN + c * 8);
    }
}

enum Epi { EPI_GELU = 0, EPI_GELU_PRE = 1, EPI_LINEAR = 2, EPI_BWD_GELU = 3 };

struct ConvArgs {
    const uint16_t* in;   // [B][N^3][CIN] bf16
    const uint16_t* w;    // [COUT][27 CIN] bf16
    const float* bias;    // [COUT] (forward modes)
    const float* pre_in;  // EPI_BWD_GELU: pre-activation at the output points [B][N^3][COUT]
    uint16_t* act_out;    // bf16 output [B][N^3][COUT] (GELU modes, EPI_BWD_GELU)
    float* f_out;         // FP32 output [B][N^3][COUT] (EPI_GELU_PRE: pre-activation; EPI_LINEAR)
    int N;
};

// grid (tiles, 1, batch); block 512.  The layer's weights are staged in LDS next to the input
// halo (row stride 27 CIN + 8 halfwords: the 16 rows of one B fragment fall in distinct banks);
// the next K step's fragments are loaded while the current step's MFMAs run; bf16 outputs leave
// through an LDS tile as 16-byte stores.  Used for the input gradient (EPI_BWD_GELU: flipped,
// transposed weights, GELU' of the stored pre-activation); the forward modes run in conv3p_kernel.
template <int CIN, int COUT, int EPI>
__global__ void __launch_bounds__(CONV_THREADS) conv3_kernel(ConvArgs a) {
    static_assert(EPI == EPI_BWD_GELU, "input gradients only");
    constexpr int SP = CIN + 8;
    constexpr int KT = 27 * CIN;
    constexpr int WS = KT + 8;
    constexpr int KSTEPS = (KT + 31) / 32;
    constexpr int NH = COUT / 16;
    constexpr int SUB = TPTS / 16 / (CONV_THREADS / 64);  // z-rows per wavefront
    constexpr int OS = COUT + 8;                          // output tile point stride (halfwords)
    static_assert(TPTS * OS <= HPTS * SP, "output tile must fit in the halo buffer");
    __shared__ __attribute__((aligned(16))) uint16_t lds[HPTS * SP + COUT * WS];
    uint16_t* ldw = lds + HPTS * SP;
    const int N = a.N;
    const long long N3 = (long long)N * N * N;
    int x0, y0, z0;
    tile_origin(blockIdx.x, N, x0, y0, z0);
    const int b = blockIdx.z;
    for (int i = threadIdx.x; i < COUT * (KT / 8); i += CONV_THREADS) {
        const int r = i / (KT / 8), c = i % (KT / 8);
        *(u32x4*)(ldw + r * WS + c * 8) = *(const u32x4*)(a.w + (long long)r * KT + c * 8);
    }
    stage_halo<CIN, SP>(lds, a.in + (long long)b * N3 * CIN, N, x0, y0, z0);
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int m = lane & 15, g = lane >> 4;
    floatx4 acc[SUB][NH];
#pragma unroll
    for (int s = 0; s < SUB; ++s)
#pragma unroll
        for (int h = 0; h < NH; ++h) acc[s][h] = floatx4{0.f, 0.f, 0.f, 0.f};
    auto load = [&](int ks, bf16x8 (&af)[SUB], bf16x8 (&bw)[NH]) {
        const int k0 = 32 * ks + 8 * g;
        const int tap = k0 / CIN, ch0 = k0 % CIN;
        if (tap < 27) {
            const int dx = tap / 9, dy = (tap / 3) % 3, dz = tap % 3;
#pragma unroll
            for (int h = 0; h < NH; ++h) bw[h] = ld16(ldw + (h * 16 + m) * WS + k0);
#pragma unroll
            for (int s = 0; s < SUB; ++s) {
                const int st = wv * SUB + s, ix = st / TYT, iy = st % TYT;
                af[s] = ld16(lds + (((ix + dx) * HY + (iy + dy)) * HZ + (m + dz)) * SP + ch0);
            }
        } else {  // K padding of CIN = 16 (27 taps = 13.5 K steps)
#pragma unroll
            for (int h = 0; h < NH; ++h) bw[h] = bf16x8{};
#pragma unroll
            for (int s = 0; s < SUB; ++s) af[s] = bf16x8{};
        }
    };
    bf16x8 af[SUB], bw[NH];
    load(0, af, bw);
    for (int ks = 0; ks < KSTEPS; ++ks) {
        bf16x8 an[SUB], bn[NH];
        if (ks + 1 < KSTEPS) load(ks + 1, an, bn);
#pragma unroll
        for (int s = 0; s < SUB; ++s)
#pragma unroll
            for (int h = 0; h < NH; ++h) acc[s][h] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[s], bw[h], acc[s][h], 0, 0, 0);
        if (ks + 1 < KSTEPS) {
#pragma unroll
            for (int s = 0; s < SUB; ++s) af[s] = an[s];
#pragma unroll
            for (int h = 0; h < NH; ++h) bw[h] = bn[h];
        }
    }
    // epilogue: lane holds rows (points) 4g..4g+3 of column (channel) m of each block
    __syncthreads();  // the halo buffer becomes the output tile
#pragma unroll
    for (int s = 0; s < SUB; ++s) {
        const int st = wv * SUB + s, ix = st / TYT, iy = st % TYT;
        const long long row = (long long)b * N3 + ((long long)(x0 + ix) * N + (y0 + iy)) * N + z0;
#pragma unroll
        for (int h = 0; h < NH; ++h) {
            const int co = h * 16 + m;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const long long idx = (row + 4 * g + i) * COUT + co;
                const int tp = st * TZT + 4 * g + i;  // point index within the tile (z fastest)
                lds[tp * OS + co] = f2bf(acc[s][h][i] * gelu_grad(a.pre_in[idx]));
            }
        }
    }
    __syncthreads();
    constexpr int CH = COUT / 8;
    for (int i = threadIdx.x; i < TPTS * CH; i += CONV_THREADS) {
        const int tp = i / CH, c = i % CH;
        const int iz = tp % TZT, iy = (tp / TZT) % TYT, ix = tp / (TZT * TYT);
        const long long gp = (long long)b * N3 + ((long long)(x0 + ix) * N + (y0 + iy)) * N + (z0 + iz);
        *(u32x4*)(a.act_out + gp * COUT + c * 8) = *(const u32x4*)(lds + tp * OS + c * 8);
    }
}

// Forward convolution (forward modes), persistent.  Blocks of 256 threads walk the (sample, tile)
// items of 4 x 4 x 16 output points, two blocks per CU.  The 6 x 6 x 18 periodic halo of an item
// goes from HBM to LDS through the global->LDS DMA (64 lanes x 16 bytes per instruction); a block
// has one halo buffer and a separate output tile, and copies the next item's halo as soon as every
// wavefront has read the current one, while its own epilogue (GELU, output tile, stores) runs; the
// two blocks of a CU drift apart, so one's epilogue and copies overlap the other's MFMAs.  The
// layer's weights stay in registers for the whole kernel: a wavefront holds 16 output channels'
// (with COUT = 32 the four wavefronts split them into two halves over 8 z-rows each, with 16 each
// holds them all over 4 z-rows), so LDS carries only the A operand.  The DMA writes each halo
// point's CIN channels contiguously (no padding), so the 16-byte chunks of a point are swizzled
// instead: chunk c of the point with halo z index hz sits in slot c ^ ((hz >> SH) & (CH - 1)),
// which puts the 8 points that one LDS cycle serves in distinct banks.  GEMM view per item as
// conv3_kernel's: M = 256 points (16 z-rows), N = COUT, K = 27 taps x CIN in steps of 32.
constexpr int PX = 4, PY = 4, PZ = 16;
constexpr int PHY = PY + 2, PHZ = PZ + 2, PHPTS = (PX + 2) * PHY * PHZ, PPTS = PX * PY * PZ;
constexpr int Q_THREADS = 256;

template <int CIN, int COUT, int EPI>
__global__ void __launch_bounds__(Q_THREADS, 2) conv3p_kernel(ConvArgs a, int ntile, int nitems) {
    static_assert(EPI == EPI_GELU || EPI == EPI_GELU_PRE || EPI == EPI_LINEAR, "forward modes only");
    constexpr int CH = CIN / 8;                  // 16-byte chunks per point
    constexpr int SH = (CH == 4) ? 1 : 2;        // swizzle shift (CH = 2 or 4)
    static_assert(CH == 2 || CH == 4, "CIN = 16 or 32");
    constexpr int KT = 27 * CIN, KSTEPS = (KT + 31) / 32, NH = COUT / 16;
    constexpr int NWAVE = Q_THREADS / 64;
    constexpr int HS = (NH >= 2) ? 2 : 1;        // channel groups of wavefronts
    constexpr int NHW = NH / HS;                 // 16-channel blocks per wavefront
    constexpr int SUB = PPTS / 16 / (NWAVE / HS);  // z-rows per wavefront
    constexpr int OS = COUT + 8;                 // output tile point stride (halfwords)
    constexpr int HB = PHPTS * CIN;              // halfwords of the halo buffer
    constexpr int NCHUNK = PHPTS * CH;
    constexpr int ROUNDS = (NCHUNK + Q_THREADS - 1) / Q_THREADS;
    constexpr int OT = (EPI != EPI_LINEAR) ? PPTS * OS : 8;  // bf16 output tile (GELU modes)
    __shared__ __attribute__((aligned(16))) uint16_t halo[HB];
    __shared__ __attribute__((aligned(16))) uint16_t outt[OT];
    const int N = a.N;
    const long long N3 = (long long)N * N * N;
    const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int m = lane & 15, g = lane >> 4;
    const int wrow = wv % (NWAVE / HS), hgrp = wv / (NWAVE / HS);  // z-rows wrow SUB .., channels hgrp NHW 16 ..

    // this thread's share of a halo copy, the same for every item: chunk f = (r NWAVE + wv) 64 + lane
    // is chunk src of halo point (hx, hy, hz), held as its byte offset from the point (x0 - 1,
    // y0 - 1, z0 - 1) without wrapping, 2 (((hx N + hy) N + hz) CIN + 8 src) (bits 0-25, N <= 256),
    // with flags for the halo faces that may wrap: hx = 0 (bit 26), hx = PX + 1 (27), hy = 0 (28),
    // hy = PY + 1 (29), hz = 0 (30), hz = PZ + 1 (31); ~0: past the halo.  Kept in LDS (each
    // thread reads back only its own entries): in registers they would spill, and a spill reload
    // waits for every copy in flight.
    const bool off32 = N <= 256;  // sample offsets fit 32 bits: 2 N^3 CIN bytes < 2^31
    constexpr int RFULL = NCHUNK / (NWAVE * 64);  // rounds with every lane inside the halo
    __shared__ unsigned hoff_t[ROUNDS * Q_THREADS];
    unsigned* hoff = hoff_t + threadIdx.x;  // entry r at hoff[r Q_THREADS]
#pragma unroll
    for (int r = 0; r < ROUNDS; ++r) {
        const int f = (r * NWAVE + wv) * 64 + lane;
        if (f < NCHUNK) {
            const int q = f / CH, slot = f % CH;
            const int hz = q % PHZ, hy = (q / PHZ) % PHY, hx = q / (PHZ * PHY);
            const int src = slot ^ ((hz >> SH) & (CH - 1));
            hoff[r * Q_THREADS] = (unsigned)((2 * (((hx * N + hy) * N + hz) * CIN + src * 8)) & 0x3ffffff) |
                      ((unsigned)(hx == 0) << 26) | ((unsigned)(hx == PX + 1) << 27) | ((unsigned)(hy == 0) << 28) |
                      ((unsigned)(hy == PY + 1) << 29) | ((unsigned)(hz == 0) << 30) | ((unsigned)(hz == PZ + 1) << 31);
        } else {
            hoff[r * Q_THREADS] = ~0u;
        }
    }
    auto origin = [&](int item, int& b, int& x0, int& y0, int& z0) {
        b = item / ntile;
        int t = item % ntile;
        const int ntz = N / PZ, nty = N / PY;
        z0 = (t % ntz) * PZ;
        t /= ntz;
        y0 = (t % nty) * PY;
        x0 = (t / nty) * PX;
    };
    auto issue = [&](int item) {
        int b, x0, y0, z0;
        origin(item, b, x0, y0, z0);
        const uint16_t* in = a.in + (long long)b * N3 * CIN;
        if (off32) {
            // byte offset = the item's corner (x0 - 1, y0 - 1, z0 - 1), possibly negative, + the
            // lane's unwrapped offset + N along each axis whose face lies outside the box for this
            // item (0 or -1 masks from the face flags); three loops, by the axes that wrap
            const int ib = 2 * ((((x0 - 1) * N + (y0 - 1)) * N + (z0 - 1)) * CIN);
            const bool wxl = x0 == 0, wxh = x0 + PX == N, wyl = y0 == 0, wyh = y0 + PY == N;
            const bool wzl = z0 == 0, wzh = z0 + PZ == N;
            const int sx = 2 * N * N * CIN, sy = 2 * N * CIN, sz = 2 * CIN;
            auto one = [&](int r, unsigned d, unsigned off) {
                if (r < RFULL || d != ~0u) dma16o(in, off, halo + (r * NWAVE + wv) * 64 * 8);
            };
            if (!(wxl || wxh || wyl || wyh || wzl || wzh)) {
#pragma unroll
                for (int r = 0; r < ROUNDS; ++r) {
                    const unsigned d = hoff[r * Q_THREADS];
                    one(r, d, (d & 0x3ffffffu) + (unsigned)ib);
                }
            } else if (!(wxl || wxh || wyl || wyh) && wzl != wzh) {
                const int bz = wzl ? 30 : 31, cz = (wzl ? N : -N) * sz;
#pragma unroll
                for (int r = 0; r < ROUNDS; ++r) {
                    const unsigned d = hoff[r * Q_THREADS];
                    one(r, d, (d & 0x3ffffffu) + (unsigned)ib + (unsigned)(__builtin_amdgcn_sbfe((int)d, bz, 1) & cz));
                }
            } else {
                const int cxl = wxl ? N * sx : 0, cxh = wxh ? -N * sx : 0, cyl = wyl ? N * sy : 0;
                const int cyh = wyh ? -N * sy : 0, czl = wzl ? N * sz : 0, czh = wzh ? -N * sz : 0;
#pragma unroll
                for (int r = 0; r < ROUNDS; ++r) {
                    const unsigned d = hoff[r * Q_THREADS];
                    const int c = (__builtin_amdgcn_sbfe((int)d, 26, 1) & cxl) + (__builtin_amdgcn_sbfe((int)d, 27, 1) & cxh) +
                                  (__builtin_amdgcn_sbfe((int)d, 28, 1) & cyl) + (__builtin_amdgcn_sbfe((int)d, 29, 1) & cyh) +
                                  (__builtin_amdgcn_sbfe((int)d, 30, 1) & czl) + (__builtin_amdgcn_sbfe((int)d, 31, 1) & czh);
                    one(r, d, (d & 0x3ffffffu) + (unsigned)ib + (unsigned)c);
                }
            }
        } else {
#pragma unroll
            for (int r = 0; r < ROUNDS; ++r) {
                const int f0 = (r * NWAVE + wv) * 64;
                int f = f0 + lane;
                asm volatile("" : "+v"(f));
                if (f < NCHUNK) {
                    const int q = f / CH, slot = f % CH;
                    const int hz = q % PHZ, hy = (q / PHZ) % PHY, hx = q / (PHZ * PHY);
                    const int src = slot ^ ((hz >> SH) & (CH - 1));
                    const int gx = wrap(x0 + hx - 1, N), gy = wrap(y0 + hy - 1, N), gz = wrap(z0 + hz - 1, N);
                    const uint16_t* gp = in + (((long long)gx * N + gy) * N + gz) * CIN + src * 8;
                    dma16((const void*)gp, halo + f0 * 8);
                }
            }
        }
    };

    // weights of K step ks: tap dx 9 + dy 3 + dz of the step's (dx, dz, dy) order for CIN = 32,
    // taps 2 ks, 2 ks + 1 for CIN = 16
    auto kmap = [](int ks) { return (CIN == 32) ? (ks / 9) * 9 + (ks % 3) * 3 + (ks / 3) % 3 : ks; };
    bf16x8 bw[KSTEPS][NHW];
#pragma unroll
    for (int ks = 0; ks < KSTEPS; ++ks) {
        const int k0 = 32 * ks + 8 * g;
#pragma unroll
        for (int h = 0; h < NHW; ++h)
            bw[ks][h] = (k0 < KT) ? ld16(a.w + ((hgrp * NHW + h) * 16 + m) * KT + k0) : bf16x8{};
    }
    float bias[NHW];
#pragma unroll
    for (int h = 0; h < NHW; ++h) bias[h] = a.bias[(hgrp * NHW + h) * 16 + m];
    // halo z part of this lane's A fragment offsets: z index m + dz, chunk g (CIN = 32) or g % 2
    // (CIN = 16: lanes g >= 2 read the second tap of a K step)
    int zoff[3];
#pragma unroll
    for (int dz = 0; dz < 3; ++dz) {
        const int cc = (CH == 4) ? g : (g & 1);
        zoff[dz] = (m + dz) * CIN + ((cc ^ (((m + dz) >> SH) & (CH - 1))) * 8);
    }

    // Before an item: this wavefront's copies of its halo have landed (and its LDS accesses are
    // done), then every wavefront's.  The previous item's global stores, issued after those copies,
    // may stay in flight: the vector-memory counter completes in issue order, so waiting until at
    // most NST operations are outstanding (the stores one item issues) covers the copies.  The
    // first item also waits for the weight and bias reads.
    constexpr int NST = ((EPI != EPI_LINEAR) ? PPTS * (COUT / 8) / Q_THREADS : 0) + ((EPI != EPI_GELU) ? SUB * NHW * 4 : 0);
    static_assert(NST < 64, "vmcnt field");
    constexpr int WAIT_NST = (NST & 15) | (7 << 4) | ((NST >> 4) << 14);
    int item = blockIdx.x;
    if (item < nitems) issue(item);
    asm volatile("s_waitcnt vmcnt(0) lgkmcnt(0)" ::: "memory");
    while (item < nitems) {
        __builtin_amdgcn_s_waitcnt(WAIT_NST);
        asm volatile("s_barrier" ::: "memory");
        floatx4 acc[SUB][NHW];
#pragma unroll
        for (int s2 = 0; s2 < SUB; ++s2)
#pragma unroll
            for (int h = 0; h < NHW; ++h) acc[s2][h] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < KSTEPS; ++ks) {
            bf16x8 af[SUB];
#pragma unroll
            for (int s2 = 0; s2 < SUB; ++s2) {
                const int st = wrow * SUB + s2, ix = st / PY, iy = st % PY;
                if constexpr (CIN == 32) {
                    // K step ks = tap kmap(ks): the three dy taps of one (dx, dz) are consecutive,
                    // so a fragment shared by neighbouring z-rows is live for three steps only
                    const int dx = ks / 9, dz = (ks / 3) % 3, dy = ks % 3;
                    af[s2] = ld16(halo + ((ix + dx) * PHY + (iy + dy)) * PHZ * CIN + zoff[dz]);
                } else {
                    const int t0 = 2 * ks, t1 = 2 * ks + 1;
                    if (t1 < 27) {
                        const int o0 = ((ix + t0 / 9) * PHY + (iy + (t0 / 3) % 3)) * PHZ * CIN + zoff[t0 % 3];
                        const int o1 = ((ix + t1 / 9) * PHY + (iy + (t1 / 3) % 3)) * PHZ * CIN + zoff[t1 % 3];
                        af[s2] = ld16(halo + (g < 2 ? o0 : o1));
                    } else {
                        const int o0 = ((ix + t0 / 9) * PHY + (iy + (t0 / 3) % 3)) * PHZ * CIN + zoff[t0 % 3];
                        af[s2] = (g < 2) ? ld16(halo + o0) : bf16x8{};
                    }
                }
            }
#pragma unroll
            for (int s2 = 0; s2 < SUB; ++s2)
#pragma unroll
                for (int h = 0; h < NHW; ++h)
                    acc[s2][h] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[s2], bw[kmap(ks)][h], acc[s2][h], 0, 0, 0);
        }
        // every wavefront is done reading the halo: the next item's copy may overwrite it
        asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
        const int nx = item + (int)gridDim.x;
        if (nx < nitems) issue(nx);
        int b, x0, y0, z0;
        origin(item, b, x0, y0, z0);
#pragma unroll
        for (int s2 = 0; s2 < SUB; ++s2) {
            const int st = wrow * SUB + s2, ix = st / PY, iy = st % PY;
            const long long row = (long long)b * N3 + ((long long)(x0 + ix) * N + (y0 + iy)) * N + z0;
#pragma unroll
            for (int h = 0; h < NHW; ++h) {
                const int co = (hgrp * NHW + h) * 16 + m;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const long long idx = (row + 4 * g + i) * COUT + co;
                    const int tp = st * PZ + 4 * g + i;
                    const float v = acc[s2][h][i] + bias[h];
                    if constexpr (EPI == EPI_GELU) {
                        outt[tp * OS + co] = f2bf_hw(gelu_as(v));
                    } else if constexpr (EPI == EPI_GELU_PRE) {
                        a.f_out[idx] = v;
                        outt[tp * OS + co] = f2bf_hw(gelu_as(v));
                    } else {
                        a.f_out[idx] = v;
                    }
                }
            }
        }
        if constexpr (EPI != EPI_LINEAR) {
            asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
            constexpr int OC = COUT / 8;
#pragma unroll
            for (int j = 0; j < PPTS * OC / Q_THREADS; ++j) {
                const int i = j * Q_THREADS + threadIdx.x;
                const int tp = i / OC, c = i % OC;
                const int iz = tp % PZ, iy = (tp / PZ) % PY, ix = tp / (PZ * PY);
                const long long gp = (long long)b * N3 + ((long long)(x0 + ix) * N + (y0 + iy)) * N + (z0 + iz);
                *(u32x4*)(a.act_out + gp * COUT + c * 8) = *(const u32x4*)(outt + tp * OS + c * 8);
            }
        }
        item = nx;
    }
}

// Weight gradients: partial[blk][co][tap * CIN + ci] = sum over this block's points p of
// dp[p][co] * act[p + tap][ci], bpartial[blk][co] = sum dp[p][co].  GEMM view: M = COUT,
// N = 27 CIN, K = points (32 per MFMA).  Persistent over (tile, sample); wavefront w owns the
// column blocks w, w + 8, ...  grid (wgrad_blocks); block 512
template <int CIN, int COUT>
__global__ void __launch_bounds__(CONV_THREADS) wgrad_kernel(const uint16_t* __restrict__ act, const uint16_t* __restrict__ dp,
                                                             float* __restrict__ partial, float* __restrict__ bpartial, int N,
                                                             int B) {
    constexpr int SP = CIN + 8, SPD = COUT + 8;
    constexpr int KT = 27 * CIN;
    constexpr int NT = KT / 16, MT = COUT / 16;
    constexpr int NTW = (NT + 7) / 8;
    __shared__ __attribute__((aligned(16))) uint16_t lds[HPTS * SP];
    __shared__ __attribute__((aligned(16))) uint16_t ldd[TPTS * SPD];
    __shared__ float bred[CONV_THREADS];
    const long long N3 = (long long)N * N * N;
    const int ntiles = (N / TXT) * (N / TYT) * (N / TZT);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int m = lane & 15, g = lane >> 4;
    floatx4 acc[NTW][MT];
#pragma unroll
    for (int r = 0; r < NTW; ++r)
#pragma unroll
        for (int q = 0; q < MT; ++q) acc[r][q] = floatx4{0.f, 0.f, 0.f, 0.f};
    float bsum = 0.f;
    const int bco = threadIdx.x % COUT;
    for (int job = blockIdx.x; job < ntiles * B; job += gridDim.x) {
        const int b = job / ntiles, t = job % ntiles;
        int x0, y0, z0;
        tile_origin(t, N, x0, y0, z0);
        __syncthreads();  // previous tile's LDS reads are done
        stage_halo<CIN, SP>(lds, act + (long long)b * N3 * CIN, N, x0, y0, z0);
        {
            constexpr int CH = COUT / 8;
            const uint16_t* src = dp + (long long)b * N3 * COUT;
            for (int i = threadIdx.x; i < TPTS * CH; i += blockDim.x) {
                const int p = i / CH, c = i % CH;
                const int iz = p % TZT, iy = (p / TZT) % TYT, ix = p / (TZT * TYT);
                const long long gp = ((long long)(x0 + ix) * N + (y0 + iy)) * N + (z0 + iz);
                *(u32x4*)(ldd + p * SPD + c * 8) = *(const u32x4*)(src + gp * COUT + c * 8);
            }
        }
        __syncthreads();
        for (int p = threadIdx.x / COUT; p < TPTS; p += CONV_THREADS / COUT) bsum += bf2f(ldd[p * SPD + bco]);
        for (int ks = 0; ks < TPTS / 32; ++ks) {
            // this lane's 8 points: 32 ks + 8 g + j lie in one z-row (z = 8 (g & 1) + j)
            const int p0 = 32 * ks + 8 * g;
            const int ix = p0 / (TZT * TYT), iy = (p0 / TZT) % TYT, iz0 = p0 % TZT;
            bf16x8 af[MT];
#pragma unroll
            for (int q = 0; q < MT; ++q) {
#pragma unroll
                for (int j = 0; j < 8; ++j) af[q][j] = __builtin_bit_cast(__bf16, ldd[(p0 + j) * SPD + q * 16 + m]);
            }
#pragma unroll
            for (int r = 0; r < NTW; ++r) {
                const int nt = wv + 8 * r;
                if (nt < NT) {
                    const int K = nt * 16 + m, tap = K / CIN, ci = K % CIN;
                    const int dx = tap / 9, dy = (tap / 3) % 3, dz = tap % 3;
                    const int qb = ((ix + dx) * HY + (iy + dy)) * HZ + (iz0 + dz);
                    bf16x8 bv;
#pragma unroll
                    for (int j = 0; j < 8; ++j) bv[j] = __builtin_bit_cast(__bf16, lds[(qb + j) * SP + ci]);
#pragma unroll
                    for (int q = 0; q < MT; ++q) acc[r][q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[q], bv, acc[r][q], 0, 0, 0);
                }
            }
        }
    }
    // D[4 g + i][m] of block (q, nt): co = 16 q + 4 g + i, K = 16 nt + m
    float* P = partial + (long long)blockIdx.x * COUT * KT;
#pragma unroll
    for (int r = 0; r < NTW; ++r) {
        const int nt = wv + 8 * r;
        if (nt < NT) {
#pragma unroll
            for (int q = 0; q < MT; ++q)
#pragma unroll
                for (int i = 0; i < 4; ++i) P[(long long)(16 * q + 4 * g + i) * KT + 16 * nt + m] = acc[r][q][i];
        }
    }
    bred[threadIdx.x] = bsum;
    __syncthreads();
    if (threadIdx.x < COUT) {
        float s = 0.f;
        for (int k = threadIdx.x; k < CONV_THREADS; k += COUT) s += bred[k];
        bpartial[(long long)blockIdx.x * COUT + threadIdx.x] = s;
    }
}

// grad[e] = sum over blocks of partial (fixed order); e < COUT*KT: weights, then COUT biases
__global__ void wgrad_reduce_kernel(const float* __restrict__ partial, const float* __restrict__ bpartial, float* __restrict__ grad,
                                    int nblk, int nw, int nb) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < nw) {
        float s = 0.f;
        for (int k = 0; k < nblk; ++k) s += partial[(long long)k * nw + e];
        grad[e] = s;
    } else if (e < nw + nb) {
        const int co = e - nw;
        float s = 0.f;
        for (int k = 0; k < nblk; ++k) s += bpartial[(long long)k * nb + co];
        grad[e] = s;
    }
}

// MSE: loss partials and dL/d(out) = 2 (out - t) / count as bf16 (16 channels, 6 used)
__global__ void loss_kernel(const float* __restrict__ out, const float* __restrict__ t, uint16_t* __restrict__ dp,
                            double* __restrict__ lpart, long long npts, double inv_count) {
    __shared__ double red[256];
    double acc = 0.0;
    for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < npts; p += (long long)gridDim.x * blockDim.x) {
#pragma unroll
        for (int c = 0; c < 16; ++c) {
            uint16_t v = 0;
            if (c < 6) {
                const float d = out[p * 16 + c] - t[p * 6 + c];
                acc += (double)d * d;
                v = f2bf((float)(2.0 * d * inv_count));
            }
            if (dp) dp[p * 16 + c] = v;
        }
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) lpart[blockIdx.x] = red[0];
}

__global__ void sum_kernel(const double* __restrict__ p, int n, double* __restrict__ out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double s = 0.0;
        for (int i = 0; i < n; ++i) s += p[i];
        *out = s;
    }
}

// Adam on layer l's master parameters (padded entries are held at zero), then the bf16 copies
struct AdamArgs {
    float *p, *m, *v;
    const float* g;
    uint16_t *wf, *wb;
    int cin, cout, cin_real, cout_real;
    float lr, b1, b2, eps, c1, c2;  // c1 = 1 / (1 - b1^t), c2 = 1 / (1 - b2^t)
    int update;                     // 0: only refresh the bf16 copies
};
__global__ void adam_kernel(AdamArgs a) {
    const long long KT = 27LL * a.cin, nw = (long long)a.cout * KT, n = nw + a.cout;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x) {
        bool valid;
        int co, K = 0;
        if (e < nw) {
            co = (int)(e / KT);
            K = (int)(e % KT);
            valid = co < a.cout_real && (K % a.cin) < a.cin_real;
        } else {
            co = (int)(e - nw);
            valid = co < a.cout_real;
        }
        float p = a.p[e];
        if (a.update) {
            if (valid) {
                const float g = a.g[e];
                const float m = a.b1 * a.m[e] + (1.f - a.b1) * g;
                const float v = a.b2 * a.v[e] + (1.f - a.b2) * g * g;
                a.m[e] = m;
                a.v[e] = v;
                p -= a.lr * (m * a.c1) / (sqrtf(v * a.c2) + a.eps);
            } else {
                p = 0.f;
            }
            a.p[e] = p;
        }
        if (e < nw) {
            const uint16_t h = f2bf(p);
            a.wf[e] = h;
            const int tap = K / a.cin, ci = K % a.cin;
            a.wb[(long long)ci * 27 * a.cout + (26 - tap) * a.cout + co] = h;
        }
    }
}

// physical [n][9] -> normalised bf16 [n][16]; physical [n][6] -> normalised FP32 [n][6]
__global__ void pack_x_kernel(const float* __restrict__ x, uint16_t* __restrict__ out, long long n, float inv) {
    for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < n; p += (long long)gridDim.x * blockDim.x)
#pragma unroll
        for (int c = 0; c < 16; ++c) out[p * 16 + c] = c < 9 ? f2bf(x[p * 9 + c] * inv) : (uint16_t)0;
}
__global__ void pack_y_kernel(const float* __restrict__ y, float* __restrict__ out, long long n, float i0, float i1, float i2,
                              float i3, float i4, float i5) {
    const float inv[6] = {i0, i1, i2, i3, i4, i5};
    for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < n; p += (long long)gridDim.x * blockDim.x)
#pragma unroll
        for (int c = 0; c < 6; ++c) out[p * 6 + c] = y[p * 6 + c] * inv[c];
}
__global__ void unpack_out_kernel(const float* __restrict__ in, float* __restrict__ y, long long n) {
    for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < n; p += (long long)gridDim.x * blockDim.x)
#pragma unroll
        for (int c = 0; c < 6; ++c) y[p * 6 + c] = in[p * 16 + c];
}

int cfail(nsg_cnn* c, int code, const char* fmt, ...) __attribute__((format(printf, 3, 4)));
int cfail(nsg_cnn* c, int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    return code;
}

#define CHK(c, expr)                                                                                           \
    do {                                                                                                       \
        hipError_t e_ = (expr);                                                                                \
        if (e_ != hipSuccess) return cfail((c), NSG_ERR_DEVICE, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
    } while (0)
#define LCHK(c) CHK(c, hipGetLastError())

int ntiles(const nsg_cnn* c) { return (c->N / TXT) * (c->N / TYT) * (c->N / TZT); }

template <int CIN, int COUT, int EPI>
int launch_conv(nsg_cnn* c, hipStream_t s, int B, const ConvArgs& a) {
    if constexpr (EPI != EPI_BWD_GELU) {
        const int nt = (c->N / PX) * (c->N / PY) * (c->N / PZ);
        const int items = nt * B;
        // two blocks per CU: one 64^3 forward took 0.071 ms against 0.081-0.087 ms with 1, 3, 4 or 6
        // blocks per CU, a batch of 8 0.355-0.360 ms against 0.37-0.51 ms
        const int blocks = std::min(items, 2 * c->ncu);
        hipLaunchKernelGGL((conv3p_kernel<CIN, COUT, EPI>), dim3(blocks), dim3(Q_THREADS), 0, s, a, nt, items);
    } else {
        hipLaunchKernelGGL((conv3_kernel<CIN, COUT, EPI>), dim3(ntiles(c), 1, B), dim3(CONV_THREADS), 0, s, a);
    }
    LCHK(c);
    return NSG_OK;
}

template <int CIN, int COUT>
int launch_wgrad(nsg_cnn* c, hipStream_t s, int B, const uint16_t* act, const uint16_t* dp, int layer) {
    const int nblk = c->wgrad_blocks;
    const int nw = COUT * 27 * CIN;
    float* bpart = c->partial + (size_t)nblk * nw;
    hipLaunchKernelGGL((wgrad_kernel<CIN, COUT>), dim3(nblk), dim3(CONV_THREADS), 0, s, act, dp, c->partial, bpart, c->N, B);
    LCHK(c);
    hipLaunchKernelGGL(wgrad_reduce_kernel, dim3((nw + COUT + 255) / 256), dim3(256), 0, s, (const float*)c->partial,
                       (const float*)bpart, c->grad + loffset(layer), nblk, nw, COUT);
    LCHK(c);
    return NSG_OK;
}

int refresh_weights(nsg_cnn* c, hipStream_t s, bool update, float lr) {
    const float b1 = 0.9f, b2 = 0.999f;
    double c1 = 1.0, c2 = 1.0;
    if (update) {
        c->adam_t += 1;
        c1 = 1.0 / (1.0 - std::pow((double)b1, (double)c->adam_t));
        c2 = 1.0 / (1.0 - std::pow((double)b2, (double)c->adam_t));
    }
    for (int l = 0; l < NLAYER; ++l) {
        AdamArgs a;
        const long long o = loffset(l);
        a.p = c->master + o; a.m = c->adam_m + o; a.v = c->adam_v + o; a.g = c->grad + o;
        a.wf = c->wf[l]; a.wb = c->wb[l];
        a.cin = CPAD[l]; a.cout = CPAD[l + 1]; a.cin_real = CREAL[l]; a.cout_real = CREAL[l + 1];
        a.lr = lr; a.b1 = b1; a.b2 = b2; a.eps = 1e-8f; a.c1 = (float)c1; a.c2 = (float)c2;
        a.update = update ? 1 : 0;
        hipLaunchKernelGGL(adam_kernel, dim3(128), dim3(256), 0, s, a);
        LCHK(c);
    }
    return NSG_OK;
}

int loss_dev(nsg_cnn* c, hipStream_t s, int B, bool want_grad, double* loss) {
    const long long npts = (long long)B * c->N3;
    const int nb = 256;
    hipLaunchKernelGGL(loss_kernel, dim3(nb), dim3(256), 0, s, (const float*)c->out, (const float*)c->target,
                       want_grad ? c->dp[0] : (uint16_t*)nullptr, c->lpart, npts, 1.0 / (6.0 * (double)npts));
    LCHK(c);
    hipLaunchKernelGGL(sum_kernel, dim3(1), dim3(64), 0, s, (const double*)c->lpart, nb, c->lpart + nb);
    LCHK(c);
    CHK(c, hipMemcpyAsync(loss, c->lpart + nb, sizeof(double), hipMemcpyDeviceToHost, s));
    CHK(c, hipStreamSynchronize(s));
    *loss /= 6.0 * (double)npts;
    return NSG_OK;
}

// backward of the whole network from dp[0] = dL/d(out) (16 ch); gradients into c->grad
int backward_dev(nsg_cnn* c, hipStream_t s, int B) {
    int rc;
    // layer 4 (32 -> 16): weights, then dL/d(pre3) = (W4^T * dp) gelu'(pre3)
    if ((rc = launch_wgrad<32, 16>(c, s, B, c->act[3], c->dp[0], 3))) return rc;
    {
        ConvArgs a{c->dp[0], c->wb[3], nullptr, c->pre[2], c->dp[1], nullptr, c->N};
        if ((rc = launch_conv<16, 32, EPI_BWD_GELU>(c, s, B, a))) return rc;
    }
    if ((rc = launch_wgrad<32, 32>(c, s, B, c->act[2], c->dp[1], 2))) return rc;
    {
        ConvArgs a{c->dp[1], c->wb[2], nullptr, c->pre[1], c->dp[0], nullptr, c->N};
        if ((rc = launch_conv<32, 32, EPI_BWD_GELU>(c, s, B, a))) return rc;
    }
    if ((rc = launch_wgrad<32, 32>(c, s, B, c->act[1], c->dp[0], 1))) return rc;
    {
        ConvArgs a{c->dp[0], c->wb[1], nullptr, c->pre[0], c->dp[1], nullptr, c->N};
        if ((rc = launch_conv<32, 32, EPI_BWD_GELU>(c, s, B, a))) return rc;
    }
    return launch_wgrad<16, 32>(c, s, B, c->act[0], c->dp[1], 0);
}

// host parameter order (torch conv3d, unpadded) <-> padded master layout
void host_to_master(const float* in, std::vector<float>& m) {
    m.assign(MASTER_COUNT, 0.f);
    long long o = 0;
    for (int l = 0; l < NLAYER; ++l) {
        const int ci_r = CREAL[l], co_r = CREAL[l + 1], cin = CPAD[l];
        float* W = m.data() + loffset(l);
        for (int co = 0; co < co_r; ++co)
            for (int ci = 0; ci < ci_r; ++ci)
                for (int tap = 0; tap < 27; ++tap) W[(long long)co * 27 * cin + tap * cin + ci] = in[o++];
        for (int co = 0; co < co_r; ++co) W[wcount(l) + co] = in[o++];
    }
}
void master_to_host(const std::vector<float>& m, float* out) {
    long long o = 0;
    for (int l = 0; l < NLAYER; ++l) {
        const int ci_r = CREAL[l], co_r = CREAL[l + 1], cin = CPAD[l];
        const float* W = m.data() + loffset(l);
        for (int co = 0; co < co_r; ++co)
            for (int ci = 0; ci < ci_r; ++ci)
                for (int tap = 0; tap < 27; ++tap) out[o++] = W[(long long)co * 27 * cin + tap * cin + ci];
        for (int co = 0; co < co_r; ++co) out[o++] = W[wcount(l) + co];
    }
}

uint64_t splitmix(uint64_t& s) {
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

void free_cnn(nsg_cnn* c) {
    void* bufs[] = {c->master, c->adam_m, c->adam_v, c->grad, c->out, c->target, c->partial, c->lpart,
                    c->dp[0], c->dp[1]};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    for (int l = 0; l < NLAYER; ++l) {
        if (c->wf[l]) (void)hipFree(c->wf[l]);
        if (c->wb[l]) (void)hipFree(c->wb[l]);
        if (c->act[l]) (void)hipFree(c->act[l]);
    }
    for (int l = 0; l < NLAYER - 1; ++l)
        if (c->pre[l]) (void)hipFree(c->pre[l]);
    for (auto* p : c->xs) (void)hipFree(p);
    for (auto* p : c->ys) (void)hipFree(p);
    for (auto& e : c->ev)
        if (e) (void)hipEventDestroy(e);
    if (c->s) (void)hipStreamDestroy(c->s);
}

int upload_batch(nsg_cnn* c, const float* x, const float* t, int B) {
    // host x [B][N^3][9] and t [B][N^3][6], already normalised
    const long long n = (long long)B * c->N3;
    float* hx = nullptr;
    CHK(c, hipMalloc(&hx, sizeof(float) * n * 9));
    hipError_t e = hipMemcpyAsync(hx, x, sizeof(float) * n * 9, hipMemcpyHostToDevice, c->s);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(pack_x_kernel, dim3(1024), dim3(256), 0, c->s, (const float*)hx, c->act[0], n, 1.0f);
        e = hipGetLastError();
    }
    if (e == hipSuccess && t) e = hipMemcpyAsync(c->target, t, sizeof(float) * n * 6, hipMemcpyHostToDevice, c->s);
    if (e == hipSuccess) e = hipStreamSynchronize(c->s);
    (void)hipFree(hx);
    if (e != hipSuccess) return cfail(c, NSG_ERR_DEVICE, "batch upload failed: %s", hipGetErrorString(e));
    return NSG_OK;
}

}  // namespace

namespace nsgcnn {
int forward_dev(nsg_cnn* c, hipStream_t s, int B, bool train) {
    int rc;
    const float* bias[NLAYER];
    for (int l = 0; l < NLAYER; ++l) bias[l] = c->master + loffset(l) + wcount(l);
    if (train) {
        ConvArgs a1{c->act[0], c->wf[0], bias[0], nullptr, c->act[1], c->pre[0], c->N};
        if ((rc = launch_conv<16, 32, EPI_GELU_PRE>(c, s, B, a1))) return rc;
        ConvArgs a2{c->act[1], c->wf[1], bias[1], nullptr, c->act[2], c->pre[1], c->N};
        if ((rc = launch_conv<32, 32, EPI_GELU_PRE>(c, s, B, a2))) return rc;
        ConvArgs a3{c->act[2], c->wf[2], bias[2], nullptr, c->act[3], c->pre[2], c->N};
        if ((rc = launch_conv<32, 32, EPI_GELU_PRE>(c, s, B, a3))) return rc;
    } else {
        ConvArgs a1{c->act[0], c->wf[0], bias[0], nullptr, c->act[1], nullptr, c->N};
        if ((rc = launch_conv<16, 32, EPI_GELU>(c, s, B, a1))) return rc;
        ConvArgs a2{c->act[1], c->wf[1], bias[1], nullptr, c->act[2], nullptr, c->N};
        if ((rc = launch_conv<32, 32, EPI_GELU>(c, s, B, a2))) return rc;
        ConvArgs a3{c->act[2], c->wf[2], bias[2], nullptr, c->act[3], nullptr, c->N};
        if ((rc = launch_conv<32, 32, EPI_GELU>(c, s, B, a3))) return rc;
    }
    ConvArgs a4{c->act[3], c->wf[3], bias[3], nullptr, nullptr, c->out, c->N};
    return launch_conv<32, 16, EPI_LINEAR>(c, s, B, a4);
}
}  // namespace nsgcnn

// ------------------------------------------------------------------------------------------
// C ABI
// ------------------------------------------------------------------------------------------
long long nsg_cnn_param_count(void) { return PARAM_COUNT; }

const char* nsg_cnn_last_error(const nsg_cnn* c) { return c ? c->err.c_str() : "null network"; }

void nsg_cnn_destroy(nsg_cnn* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->s) (void)hipStreamSynchronize(c->s);
    free_cnn(c);
    delete c;
}

int nsg_cnn_create(int N, int max_batch, uint64_t seed, int device, nsg_cnn** out) {
    if (!out) return NSG_ERR_INVALID;
    *out = nullptr;
    nsg_cnn* c = new (std::nothrow) nsg_cnn();
    if (!c) return NSG_ERR_OOM;
    *out = c;
    if (N < 16 || N % 16 != 0) return cfail(c, NSG_ERR_INVALID, "grid N=%d must be a multiple of 16", N);
    if (max_batch < 1) return cfail(c, NSG_ERR_INVALID, "max_batch must be >= 1");
    c->device = device;
    c->N = N;
    c->maxB = max_batch;
    c->N3 = (long long)N * N * N;
    CHK(c, hipSetDevice(device));
    CHK(c, hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking));
    CHK(c, hipEventCreate(&c->ev[0]));
    CHK(c, hipEventCreate(&c->ev[1]));
    {
        hipDeviceProp_t prop;
        CHK(c, hipGetDeviceProperties(&prop, device));
        c->wgrad_blocks = std::max(64, prop.multiProcessorCount);
        c->ncu = std::max(1, prop.multiProcessorCount);
    }
    const long long np = (long long)max_batch * c->N3;
    auto alloc = [&](void** p, size_t bytes) -> int {
        hipError_t e = hipMalloc(p, bytes);
        if (e != hipSuccess) return cfail(c, NSG_ERR_OOM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
        return (int)NSG_OK;
    };
    int rc;
    for (float** p : {&c->master, &c->adam_m, &c->adam_v, &c->grad})
        if ((rc = alloc((void**)p, sizeof(float) * MASTER_COUNT))) return rc;
    for (int l = 0; l < NLAYER; ++l) {
        if ((rc = alloc((void**)&c->wf[l], 2 * wcount(l)))) return rc;
        if ((rc = alloc((void**)&c->wb[l], 2 * wcount(l)))) return rc;
        if ((rc = alloc((void**)&c->act[l], 2 * np * CPAD[l]))) return rc;
    }
    for (int l = 0; l < NLAYER - 1; ++l)
        if ((rc = alloc((void**)&c->pre[l], 4 * np * 32))) return rc;
    if ((rc = alloc((void**)&c->out, 4 * np * 16))) return rc;
    if ((rc = alloc((void**)&c->target, 4 * np * 6))) return rc;
    for (int k = 0; k < 2; ++k)
        if ((rc = alloc((void**)&c->dp[k], 2 * np * 32))) return rc;
    c->partial_n = (size_t)c->wgrad_blocks * (32 * 27 * 32 + 32);
    if ((rc = alloc((void**)&c->partial, sizeof(float) * c->partial_n))) return rc;
    if ((rc = alloc((void**)&c->lpart, sizeof(double) * 512))) return rc;
    CHK(c, hipMemset(c->adam_m, 0, sizeof(float) * MASTER_COUNT));
    CHK(c, hipMemset(c->adam_v, 0, sizeof(float) * MASTER_COUNT));
    CHK(c, hipMemset(c->grad, 0, sizeof(float) * MASTER_COUNT));
    CHK(c, hipMemset(c->act[0], 0, 2 * np * 16));
    CHK(c, hipDeviceSynchronize());
    // torch's default conv initialisation: U(-1/sqrt(fan_in), 1/sqrt(fan_in)) for weights and biases
    std::vector<float> h(PARAM_COUNT);
    uint64_t st = seed;
    long long o = 0;
    for (int l = 0; l < NLAYER; ++l) {
        const double bound = 1.0 / std::sqrt(27.0 * CREAL[l]);
        for (long long k = 0; k < rcount(l); ++k) {
            const double u = (double)(splitmix(st) >> 11) * (1.0 / 9007199254740992.0);
            h[o++] = (float)((2.0 * u - 1.0) * bound);
        }
    }
    return nsg_cnn_set_params(c, h.data());
}

int nsg_cnn_set_params(nsg_cnn* c, const float* params) {
    if (!c || !params) return NSG_ERR_INVALID;
    CHK(c, hipSetDevice(c->device));
    std::vector<float> m;
    host_to_master(params, m);
    CHK(c, hipMemcpyAsync(c->master, m.data(), sizeof(float) * MASTER_COUNT, hipMemcpyHostToDevice, c->s));
    int rc = refresh_weights(c, c->s, false, 0.f);
    if (rc) return rc;
    CHK(c, hipStreamSynchronize(c->s));
    return NSG_OK;
}

int nsg_cnn_get_params(nsg_cnn* c, float* params) {
    if (!c || !params) return NSG_ERR_INVALID;
    CHK(c, hipSetDevice(c->device));
    std::vector<float> m(MASTER_COUNT);
    CHK(c, hipMemcpyAsync(m.data(), c->master, sizeof(float) * MASTER_COUNT, hipMemcpyDeviceToHost, c->s));
    CHK(c, hipStreamSynchronize(c->s));
    master_to_host(m, params);
    return NSG_OK;
}

int nsg_cnn_set_norm(nsg_cnn* c, double sigma_in, const double* sigma_out) {
    if (!c || !(sigma_in > 0.0) || !sigma_out) return NSG_ERR_INVALID;
    for (int k = 0; k < 6; ++k)
        if (!(sigma_out[k] > 0.0)) return cfail(c, NSG_ERR_INVALID, "sigma_out[%d] must be > 0", k);
    if (!c->xs.empty()) return cfail(c, NSG_ERR_INVALID, "set the normalisation before adding samples");
    c->sig_in = sigma_in;
    for (int k = 0; k < 6; ++k) c->sig_out[k] = sigma_out[k];
    return NSG_OK;
}

int nsg_cnn_get_norm(const nsg_cnn* c, double* sigma_in, double* sigma_out) {
    if (!c) return NSG_ERR_INVALID;
    if (sigma_in) *sigma_in = c->sig_in;
    if (sigma_out)
        for (int k = 0; k < 6; ++k) sigma_out[k] = c->sig_out[k];
    return NSG_OK;
}

int nsg_cnn_forward(nsg_cnn* c, const float* x, int B, float* y) {
    if (!c || !x || !y || B < 1 || B > c->maxB) return NSG_ERR_INVALID;
    CHK(c, hipSetDevice(c->device));
    int rc = upload_batch(c, x, nullptr, B);
    if (rc) return rc;
    if ((rc = forward_dev(c, c->s, B, false))) return rc;
    const long long n = (long long)B * c->N3;
    float* hy = nullptr;
    CHK(c, hipMalloc(&hy, sizeof(float) * n * 6));
    hipLaunchKernelGGL(unpack_out_kernel, dim3(1024), dim3(256), 0, c->s, (const float*)c->out, hy, n);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(y, hy, sizeof(float) * n * 6, hipMemcpyDeviceToHost, c->s);
    if (e == hipSuccess) e = hipStreamSynchronize(c->s);
    (void)hipFree(hy);
    if (e != hipSuccess) return cfail(c, NSG_ERR_DEVICE, "forward failed: %s", hipGetErrorString(e));
    return NSG_OK;
}

int nsg_cnn_loss_grad(nsg_cnn* c, const float* x, const float* t, int B, double* loss, float* grads) {
    if (!c || !x || !t || !loss || B < 1 || B > c->maxB) return NSG_ERR_INVALID;
    CHK(c, hipSetDevice(c->device));
    int rc = upload_batch(c, x, t, B);
    if (rc) return rc;
    if ((rc = forward_dev(c, c->s, B, true))) return rc;
    if ((rc = loss_dev(c, c->s, B, grads != nullptr, loss))) return rc;
    if (!grads) return NSG_OK;
    if ((rc = backward_dev(c, c->s, B))) return rc;
    std::vector<float> m(MASTER_COUNT);
    CHK(c, hipMemcpyAsync(m.data(), c->grad, sizeof(float) * MASTER_COUNT, hipMemcpyDeviceToHost, c->s));
    CHK(c, hipStreamSynchronize(c->s));
    master_to_host(m, grads);
    return NSG_OK;
}

int nsg_cnn_add_sample(nsg_cnn* c, const float* x, const float* tau) {
    if (!c || !x || !tau) return NSG_ERR_INVALID;
    CHK(c, hipSetDevice(c->device));
    uint16_t* dx = nullptr;
    float* dy = nullptr;
    float* tmp = nullptr;
    const long long n = c->N3;
    CHK(c, hipMalloc(&dx, 2 * n * 16));
    c->xs.push_back(dx);
    CHK(c, hipMalloc(&dy, 4 * n * 6));
    c->ys.push_back(dy);
    CHK(c, hipMalloc(&tmp, 4 * n * 9));
    hipError_t e = hipMemcpyAsync(tmp, x, 4 * n * 9, hipMemcpyHostToDevice, c->s);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(pack_x_kernel, dim3(1024), dim3(256), 0, c->s, (const float*)tmp, dx, n, (float)(1.0 / c->sig_in));
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(tmp, tau, 4 * n * 6, hipMemcpyHostToDevice, c->s);
    if (e == hipSuccess) {
        const double* so = c->sig_out;
        hipLaunchKernelGGL(pack_y_kernel, dim3(1024), dim3(256), 0, c->s, (const float*)tmp, dy, n, (float)(1 / so[0]),
                           (float)(1 / so[1]), (float)(1 / so[2]), (float)(1 / so[3]), (float)(1 / so[4]), (float)(1 / so[5]));
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->s);
    (void)hipFree(tmp);
    if (e != hipSuccess) return cfail(c, NSG_ERR_DEVICE, "sample upload failed: %s", hipGetErrorString(e));
    return NSG_OK;
}

int nsg_cnn_sample_count(const nsg_cnn* c) { return c ? (int)c->xs.size() : 0; }

static int gather_batch(nsg_cnn* c, const int* idx, int B) {
    for (int b = 0; b < B; ++b) {
        CHK(c, hipMemcpyAsync(c->act[0] + (size_t)b * c->N3 * 16, c->xs[idx[b]], 2 * c->N3 * 16, hipMemcpyDeviceToDevice, c->s));
        CHK(c, hipMemcpyAsync(c->target + (size_t)b * c->N3 * 6, c->ys[idx[b]], 4 * c->N3 * 6, hipMemcpyDeviceToDevice, c->s));
    }
    return NSG_OK;
}

int nsg_cnn_train(nsg_cnn* c, int first, int count, int steps, int batch, double lr, uint64_t seed, double* loss_per_step) {
    if (!c || steps < 0 || batch < 1 || batch > c->maxB || !(lr > 0.0)) return NSG_ERR_INVALID;
    const int S = count;
    if (first < 0 || count < 1 || first + count > (int)c->xs.size())
        return cfail(c, NSG_ERR_INVALID, "sample range [%d, %d) outside the %d samples held", first, first + count, (int)c->xs.size());
    CHK(c, hipSetDevice(c->device));
    uint64_t st = seed;
    std::vector<int> idx(batch);
    for (int it = 0; it < steps; ++it) {
        for (int b = 0; b < batch; ++b) idx[b] = first + (int)(splitmix(st) % (uint64_t)S);
        int rc = gather_batch(c, idx.data(), batch);
        if (rc) return rc;
        if ((rc = forward_dev(c, c->s, batch, true))) return rc;
        double loss = 0.0;
        if ((rc = loss_dev(c, c->s, batch, true, &loss))) return rc;
        if (!std::isfinite(loss)) return cfail(c, NSG_ERR_NONFINITE, "training loss is not finite at step %d", it);
        if (loss_per_step) loss_per_step[it] = loss;
        if ((rc = backward_dev(c, c->s, batch))) return rc;
        if ((rc = refresh_weights(c, c->s, true, (float)lr))) return rc;
    }
    CHK(c, hipStreamSynchronize(c->s));
    return NSG_OK;
}

int nsg_cnn_eval(nsg_cnn* c, int first, int count, double* loss) {
    if (!c || !loss || first < 0 || count < 1 || first + count > (int)c->xs.size()) return NSG_ERR_INVALID;
    CHK(c, hipSetDevice(c->device));
    double tot = 0.0;
    for (int k = first; k < first + count; ++k) {
        int rc = gather_batch(c, &k, 1);
        if (rc) return rc;
        if ((rc = forward_dev(c, c->s, 1, false))) return rc;
        double l = 0.0;
        if ((rc = loss_dev(c, c->s, 1, false, &l))) return rc;
        tot += l;
    }
    *loss = tot / count;
    return NSG_OK;
}

int nsg_cnn_time_forward(nsg_cnn* c, int B, int iters, double* ms) {
    if (!c || !ms || B < 1 || B > c->maxB || iters < 1) return NSG_ERR_INVALID;
    CHK(c, hipSetDevice(c->device));
    int rc = forward_dev(c, c->s, B, false);
    if (rc) return rc;
    CHK(c, hipEventRecord(c->ev[0], c->s));
    for (int i = 0; i < iters; ++i)
        if ((rc = forward_dev(c, c->s, B, false))) return rc;
    CHK(c, hipEventRecord(c->ev[1], c->s));
    CHK(c, hipEventSynchronize(c->ev[1]));
    float f = 0.f;
    CHK(c, hipEventElapsedTime(&f, c->ev[0], c->ev[1]));
    *ms = f / iters;
    return NSG_OK;
}
