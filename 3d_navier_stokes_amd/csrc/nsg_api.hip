// nsg_api.hip -- C ABI implementation (include/nsg.h): batches of boxes, launch sequencing,
// transfers, the ensemble (several batches stepped by one HIP graph) and the schedule autotuner.
//
// A context holds a batch of nbox periodic boxes of one size N and precision (nsg_create makes a
// batch of one).  One RK4 step of a batch = 4 stages x 5 passes (see nsg_kernels.hpp), each pass
// one launch over every box of the batch, plus one-thread bookkeeping kernels; nothing is copied
// to the host inside a step.  The reference counterpart is SpectralSolve's loop body
// (Solver/src/solver.c:125-194) calling RK4Step (:505) -> NonlinearRHSBatch (:620).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <type_traits>
#include <vector>

#include "../../include/nsg.h"
#include "nsg_kernels.hpp"
#include "nsg_internal.hpp"

using namespace nsg;

// ------------------------------------------------------------------------------------------
// Launch schedules = FFT schedule x persistent z-pass grid.
// FFT schedules.  EX/EY/EZ: elements per thread of the x, y (length N) and z (length N/2)
// transforms, which fixes the radix sequence (nsg_fft.hpp); WX/WY: kz columns per block of the x
// and y passes.  FFT schedule 0 is the default; the sizes up to 256 have two alternatives:
// 1 = the other radices (shorter, or longer where schedule 0 is short), 2 = wider column tiles.
// 512 and 1024 keep their per-pass layouts (pass 5's LDS-pipelined reads need 128-byte rows, and
// the alternatives measured there are the ones the comments below record).
// z grids: the persistent z pass runs zgrid_of(g) blocks per CU (capped by its work); the grid
// changes no result bit (every line is transformed the same way whichever wavefront takes it, and
// the per-box max is exact).  Schedule s = g * (FFT schedules of N) + f; nsg_autotune times all.
// ------------------------------------------------------------------------------------------
struct SchedV {
    int ex64, ex32, ey64, ey32, ez64, ez32, wx64, wx32, wy64, wy32;
};
constexpr int nzp_of(int N) { return ((N / 3 + 1 + 15) / 16) * 16; }  // = Ext<N>::NZP
// 1024: pass 1 takes 8 elements per thread in 4 (FP64) or 8 (FP32) columns, 64-byte row segments
// whose two halves go to blocks launched 8 apart (xgrid); with 16 elements in 8 / 16 columns it
// spilled 630-810 bytes per lane (measured 1024^3 step: FP64 4.18 -> 4.63, FP32-bulk 6.75 -> 8.84
// steps/s with pass 5 as set by E5of / W5of).  512 FP32: pass 1 with 16 elements per thread in
// 512-thread blocks (spill-free; the 1024-thread form spills 12 bytes) ran 0.715 against 0.644 ms.
// Up to 96 this table holds the long radices (one thread per line at 8 and 12, 8 or 12 elements
// per thread above); sched0 shortens them where that measured faster.
constexpr SchedV sched_long(int N) {
    //                     ex64 ex32 ey64 ey32 ez64 ez32 wx64 wx32 wy64 wy32
    return N == 8     ? SchedV{8, 8, 8, 8, 4, 4, 8, 16, 8, 16}
           : N == 12  ? SchedV{12, 12, 12, 12, 6, 6, 8, 16, 8, 16}
           : N == 16  ? SchedV{8, 8, 8, 8, 4, 8, 8, 16, 8, 16}
           : N == 24  ? SchedV{12, 12, 12, 12, 12, 12, 8, 16, 8, 16}
           : N == 32  ? SchedV{8, 8, 8, 8, 8, 8, 8, 16, 8, 16}
           : N == 48  ? SchedV{12, 12, 12, 12, 12, 12, 8, 16, 8, 16}
           : N == 64  ? SchedV{8, 8, 8, 8, 8, 8, 8, 16, 8, 16}
           : N == 96  ? SchedV{12, 12, 12, 12, 12, 12, 8, 16, 8, 16}
           : N == 128 ? SchedV{8, 8, 8, 8, 8, 8, 8, 16, 8, 16}
           : N == 256 ? SchedV{8, 8, 8, 8, 4, 8, 8, 16, 8, 16}
           : N == 512 ? SchedV{8, 8, 8, 16, 4, 8, 8, 16, 8, 16}
                      : SchedV{8, 8, 16, 16, 8, 8, 4, 8, 8, 16};
}
constexpr int short_ept(int N) { return (N % 3 == 0) ? 6 : 4; }
// Where schedule 0 takes the short radices (4 or 6 elements per thread in the x and y passes).
// One-box steps (autotuner timings) with them against the long ones: FP64 and FP32 8^3 0.083 vs
// 0.088 ms, 12^3 0.118 vs 0.140, 16^3 0.089 vs 0.108, 24^3 0.135 vs 0.170; FP64 32^3 0.107 vs
// 0.115, 48^3 0.169 vs 0.200, 64^3 0.165 vs 0.171, 96^3 0.345 vs 0.370; FP32 48^3 0.154 vs 0.160,
// 64^3 0.132 vs 0.136, 96^3 0.270 vs 0.277, but 32^3 0.115 vs 0.108 (long kept); 128^3 and 256^3
// were faster long in both precisions.
constexpr bool short64(int N) { return N <= 96; }
constexpr bool short32(int N) { return N <= 24 || (N >= 48 && N <= 96); }
constexpr SchedV sched0(int N) {
    SchedV v = sched_long(N);
    if (short64(N)) v.ex64 = v.ey64 = short_ept(N);
    if (short32(N)) v.ex32 = v.ey32 = short_ept(N);
    return v;
}
constexpr int nsched(int N) { return N <= 256 ? 3 : 1; }  // FFT schedules
constexpr int kZGrid[] = {8, 4, 12};
constexpr int kNZGrid = (int)(sizeof(kZGrid) / sizeof(kZGrid[0]));
// z grid g of a size and precision: 128^3 FP64 (two resident blocks per CU, 2048 blocks of work)
// defaults to 4 blocks per CU, so that each block takes two line groups: 0.400 against 0.446 ms
// per step with 8 (autotuner timings, one forced box); every other size keeps 8 (at 512^3 the z
// pass ran within 0.5 % of 8 with 6, 12 or 16 blocks per CU, and 1-9 % slower with 2, 3 or 4)
template <typename T, int N> static constexpr int zgrid_of(int g) {
    return (N == 128 && std::is_same<T, double>::value && g < 2) ? kZGrid[1 - g] : kZGrid[g];
}
constexpr SchedV sched_of(int N, int S) {
    if (S == 1) {
        // the other radices: the long ones where schedule 0 is short, the short ones elsewhere
        SchedV v = sched_long(N);
        const int e = short_ept(N);
        if (!short64(N)) v.ex64 = v.ey64 = v.ez64 = e;
        if (!short32(N)) v.ex32 = v.ey32 = v.ez32 = e;
        return v;
    }
    SchedV v = sched0(N);
    if (S == 2) {
        v.wx64 = v.wy64 = 16;
        v.wx32 = v.wy32 = (nzp_of(N) % 32 == 0) ? 32 : 16;
    }
    return v;
}
template <typename T> static constexpr bool is64() { return std::is_same<T, double>::value; }
template <typename T, int N, int S> static constexpr int EXof() { return is64<T>() ? sched_of(N, S).ex64 : sched_of(N, S).ex32; }
template <typename T, int N, int S> static constexpr int EYof() { return is64<T>() ? sched_of(N, S).ey64 : sched_of(N, S).ey32; }
template <typename T, int N, int S> static constexpr int EZof() { return is64<T>() ? sched_of(N, S).ez64 : sched_of(N, S).ez32; }
template <typename T, int N, int S> static constexpr int WXof() { return is64<T>() ? sched_of(N, S).wx64 : sched_of(N, S).wx32; }
template <typename T, int N, int S> static constexpr int WYof() { return is64<T>() ? sched_of(N, S).wy64 : sched_of(N, S).wy32; }
// pass 5 holds three fields plus the epilogue operands: at 512^3 FP32 a 1024-thread block of 8
// elements per thread spills under its 128-VGPR cap; 512 threads of 16 elements do not (measured
// 0.51 -> 0.48 ms per launch).  At 1024^3 FP64 pass 5 keeps 16 elements in 8 columns: it spills, but
// ran 7.29 ms per launch against 7.69 ms for 8 elements in 4 columns (no spill).
template <typename T, int N, int S> static constexpr int E5of() {
    return ((!is64<T>() && N >= 512) || (is64<T>() && N == 1024)) ? 16 : EXof<T, N, S>();
}
template <typename T, int N, int S> static constexpr int W5of() { return (is64<T>() && N == 1024) ? 8 : WXof<T, N, S>(); }
// pass 5 of the Dormand-Prince family (RK5-CN, DP5: the final update reads five or six slopes): at
// 512^3 FP32 takes 8 elements in 8 columns (512 threads, 180 VGPRs, paired 64-byte row segments;
// the RK4 layout spilled 184 bytes per lane): RK5-CN 52.2 -> 53.1 steps/s.  FP64 keeps the RK4
// layout, which spills 84 bytes but ran faster than the spill-free 8 elements in 4 columns (RK5-CN
// 30.5 vs 29.6 steps/s)
template <typename T, int N, int S> static constexpr int E5dp() { return (N == 512 && !is64<T>()) ? 8 : E5of<T, N, S>(); }
template <typename T, int N, int S> static constexpr int W5dp() { return (N == 512 && !is64<T>()) ? 8 : W5of<T, N, S>(); }
// pass 5 with its field reads pipelined through LDS (pass5p_kernel) for the RK4-family stages at
// 512^3: measured per launch FP32 0.334 -> 0.277 ms (stages 1-3) and 0.903 -> 0.865 ms (stage 4),
// FP64 0.529 -> 0.494 and 1.296 -> 1.282 ms
template <typename T, int N, int W> static constexpr bool p5_pipelined() {
    return N == 512 && W * sizeof(cpx_t<T>) == 128;
}

template <typename F> static int dispatch_N(int N, F&& f) {
    switch (N) {
        case 8: return f(std::integral_constant<int, 8>{});
        case 12: return f(std::integral_constant<int, 12>{});
        case 16: return f(std::integral_constant<int, 16>{});
        case 24: return f(std::integral_constant<int, 24>{});
        case 32: return f(std::integral_constant<int, 32>{});
        case 48: return f(std::integral_constant<int, 48>{});
        case 64: return f(std::integral_constant<int, 64>{});
        case 96: return f(std::integral_constant<int, 96>{});
        case 128: return f(std::integral_constant<int, 128>{});
        case 256: return f(std::integral_constant<int, 256>{});
        case 512: return f(std::integral_constant<int, 512>{});
        case 1024: return f(std::integral_constant<int, 1024>{});
        default: return NSG_ERR_INVALID;
    }
}
// f(N as a constant, schedule as a constant)
template <typename F> static int dispatch_NS(int N, int S, F&& f) {
    return dispatch_N(N, [&](auto n) {
        constexpr int NN = decltype(n)::value;
        if constexpr (nsched(NN) > 1) {
            if (S == 1) return f(n, std::integral_constant<int, 1>{});
            if (S == 2) return f(n, std::integral_constant<int, 2>{});
        }
        return f(n, std::integral_constant<int, 0>{});
    });
}

// FFT schedule of a context's launch schedule
static int fft_sched(const nsg_ctx* c) {
    const int nf = dispatch_N(c->N, [](auto nn) { return nsched(decltype(nn)::value); });
    return nf > 0 ? c->sched % nf : 0;
}

// cos/sin(2 pi e / n) with octant reduction (exact at multiples of n/8)
static void twiddle_cs(long e, long n, double* c, double* s) {
    e %= n;
    if (e < 0) e += n;
    const long q = (4 * e) / n, r = 4 * e - q * n;
    double cr, sr;
    if (2 * r <= n) {
        const double a = (M_PI / 2.0) * (double)r / (double)n;
        cr = std::cos(a); sr = std::sin(a);
    } else {
        const double a = (M_PI / 2.0) * (double)(n - r) / (double)n;
        cr = std::sin(a); sr = std::cos(a);
    }
    switch (q) {
        case 0: *c = cr; *s = sr; break;
        case 1: *c = -sr; *s = cr; break;
        case 2: *c = -cr; *s = -sr; break;
        default: *c = sr; *s = -cr; break;
    }
}

static int upload_table(nsg_ctx* c, int L, double2** d64, float2** d32) {
    std::vector<double2> h64(L);
    std::vector<float2> h32(L);
    for (int e = 0; e < L; ++e) {
        double cs, sn;
        twiddle_cs(e, L, &cs, &sn);
        h64[e] = make_double2(cs, -sn);  // exp(-2 pi i e / L)
        h32[e] = make_float2((float)cs, (float)-sn);
    }
    HIPCHK(c, hipMalloc(d64, sizeof(double2) * L));
    HIPCHK(c, hipMalloc(d32, sizeof(float2) * L));
    HIPCHK(c, hipMemcpy(*d64, h64.data(), sizeof(double2) * L, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(*d32, h32.data(), sizeof(float2) * L, hipMemcpyHostToDevice));
    return NSG_OK;
}

// ------------------------------------------------------------------------------------------
// profiling hooks: kernel class k in [0,5)
// ------------------------------------------------------------------------------------------
static inline void prof_begin(nsg_ctx* c) {
    if (c->prof) (void)hipEventRecord(c->ev[0], c->s);
}
static inline void prof_end(nsg_ctx* c, int k) {
    if (!c->prof) return;
    (void)hipEventRecord(c->ev[1], c->s);
    (void)hipEventSynchronize(c->ev[1]);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, c->ev[0], c->ev[1]);
    c->prof_ms[k] += ms;
    c->prof_cnt[k] += 1;
}

// ------------------------------------------------------------------------------------------
// pass launchers over boxes [b0, b0 + nb) of a batch
// ------------------------------------------------------------------------------------------
template <typename T> static const void* twN_of(nsg_ctx* c) {
    if constexpr (is64<T>()) return c->twN64; else return c->twN32;
}
template <typename T> static const void* twM_of(nsg_ctx* c) {
    if constexpr (is64<T>()) return c->twM64; else return c->twM32;
}
template <typename T> static constexpr size_t esz() { return sizeof(cpx_t<T>); }

// grid of the x passes: (tiles, rows, boxes), or XCD-paired tiles when a tile row is under 128 B
// x-pass grid; below 128 bytes per row segment, tiles go in pairs (ntiles = NZP / W is even: NZP is
// a multiple of 16 and W <= 8 there)
template <typename T, int W> static dim3 xgrid(int ntiles, int nrows, int nb) {
    static_assert(W * sizeof(cpx_t<T>) >= 128 || 16 % W == 0, "paired tiles need an even tile count");
    if (W * sizeof(cpx_t<T>) >= 128) return dim3(ntiles, nrows, nb);
    const long long pairs = (long long)(ntiles / 2) * nrows;
    return dim3((unsigned)(16 * ((pairs + 7) / 8)), 1, nb);
}

// Dormand-Prince 5(4) tableau, Solver/src/solver.c:34-40; row s-1 forms the input of stage s.
// A32 is the published 9/40: the reference's 0.5 breaks sum_j A_3j = C3 and makes the scheme
// first order.
static const double kDPA[7][MAXSLOPES] = {
    {0, 0, 0, 0, 0, 0},
    {1.0 / 5.0, 0, 0, 0, 0, 0},
    {3.0 / 40.0, 9.0 / 40.0, 0, 0, 0, 0},
    {44.0 / 45.0, -56.0 / 15.0, 32.0 / 9.0, 0, 0, 0},
    {19372.0 / 6561.0, -25360.0 / 2187.0, 64448.0 / 6561.0, -212.0 / 729.0, 0, 0},
    {9017.0 / 3168.0, -355.0 / 33.0, 46732.0 / 5247.0, 49.0 / 176.0, -5103.0 / 18656.0, 0},
    {35.0 / 384.0, 0, 500.0 / 1113.0, 125.0 / 192.0, -2187.0 / 6784.0, 11.0 / 84.0}};

static bool dp_family(const nsg_ctx* c) { return c->p.scheme == NSG_SCHEME_RK5CN || c->p.scheme == NSG_SCHEME_DP5; }

// slopes of boxes from b0 and the tableau row of `stage` (zero row outside the Dormand-Prince family)
template <typename T> static void fill_row(const nsg_ctx* c, StageRow& r, int b0, int stage) {
    for (int j = 0; j < MAXSLOPES; ++j) {
        r.k[j] = c->K[j] ? (const char*)c->K[j] + b0 * c->sbox * esz<T>() : nullptr;
        r.a[j] = (dp_family(c) && stage >= 1 && stage <= 7 && j < c->nslopes) ? kDPA[stage - 1][j] : 0.0;
    }
}

template <typename T, int N, int S>
static int launch_pass1(nsg_ctx* c, const double2* uh, int b0, int nb, const void* kprev, int stage, int sel, int force) {
    constexpr int EPT = EXof<T, N, S>(), W = WXof<T, N, S>(), TPF = N / EPT;
    static_assert(TPF * W <= 1024, "block too large");
    P1Args a;
    a.uh = uh + b0 * c->sbox;
    a.kprev = kprev ? (const char*)kprev + b0 * c->sbox * esz<T>() : nullptr;
    fill_row<T>(c, a.row, b0, stage);
    a.out = (char*)c->B1 + b0 * c->a1box * esz<T>();
    a.twN = twN_of<T>(c); a.sc = c->sc + b0; a.fslots = c->fslots + (long long)b0 * c->NK;
    a.kf2 = c->p.force_kf * c->p.force_kf;
    a.sbox = c->sbox; a.obox = c->a1box;
    a.visc_pow = c->p.visc_pow; a.scheme = c->p.scheme; a.stage = stage; a.force = force; a.sel = sel;
    a.kmax = c->kmax; a.NK = c->NK; a.NZP = c->NZP;
    const dim3 grid = xgrid<T, W>(c->NZP / W, c->NK, nb);
    prof_begin(c);
    if (dp_family(c))
        hipLaunchKernelGGL((pass1_kernel<T, N, EPT, W, 1>), grid, dim3(TPF * W), 0, c->s, a);
    else
        hipLaunchKernelGGL((pass1_kernel<T, N, EPT, W, 0>), grid, dim3(TPF * W), 0, c->s, a);
    LAUNCHCHK(c);
    prof_end(c, 0);
    return NSG_OK;
}

// DIR > 0: A1 (B1) -> A2 (B2) on F fields;  DIR < 0: A3 (B1) -> A4 (B2) on 3 fields
template <typename T, int N, int S, int DIR>
static int launch_passy(nsg_ctx* c, int b0, int nb, int F) {
    constexpr int EPT = EYof<T, N, S>(), W = WYof<T, N, S>(), TPF = N / EPT;
    static_assert(TPF * W <= 1024, "block too large");
    PYArgs a;
    a.ibox = DIR > 0 ? c->a1box : c->a3box;
    a.obox = DIR > 0 ? c->a2box : c->a4box;
    a.in = (const char*)c->B1 + b0 * a.ibox * esz<T>();
    a.out = (char*)c->B2 + b0 * a.obox * esz<T>();
    a.F = F;
    a.twN = twN_of<T>(c); a.kmax = c->kmax; a.NK = c->NK; a.NZP = c->NZP;
    prof_begin(c);
    hipLaunchKernelGGL((passy_kernel<T, N, EPT, W, DIR>), dim3(c->NZP / W, N, F * nb), dim3(TPF * W), 0, c->s, a);
    LAUNCHCHK(c);
    prof_end(c, DIR > 0 ? 1 : 3);
    return NSG_OK;
}

// RHS z pass: persistent wavefronts when a line fits in a wavefront and N/2 is a power of two
template <typename T, int N, int S> static constexpr bool zrhs_persistent() {
    constexpr int TPFZ = N / 2 / EZof<T, N, S>();
    return TPFZ <= 64 && 64 % TPFZ == 0 && ((N / 2) & (N / 2 - 1)) == 0;
}

// Full-length paired z pass (passz_full_kernel): elements per thread of its N-point transforms, or
// 0 where the half-length form (passz_rhs_kernel) is kept.  Measured at 512^3 (isolated launches,
// same inputs): FP64 1.33 ms against 1.65 ms for the half-length form (bound by its 27 LDS
// exchanges per line), FP32 0.72 ms against 0.84 ms.  The sizes of the ensembles (32..128, and 12,
// 24, 48, 96 with a radix-3 stage, which had only the block-synchronous half-length form) use it
// too.  1024 keeps the half-length form (16 elements per thread would not fit the registers).
template <typename T, int N> static constexpr int zfull_ept() {
    if constexpr (N == 1024) return 0;
    if constexpr (is64<T>()) return (N % 3 == 0) ? 6 : (N == 512) ? 8 : 4;
    else return (N == 48 || N == 96) ? 12 : (N % 3 == 0) ? 6 : (N == 8 || N == 16) ? 4 : 8;
}
// waves per SIMD the full form is compiled for (no scratch at these bounds except FP64 96 and 512:
// 20 and 60 bytes per lane).  FP32 48 and 96 with 6 elements per thread at 3 waves per SIMD
// measured slower: 32-box 48^3 batch 0.574 -> 0.581 ms per step, 8-box 96^3 batch 0.970 -> 1.004 ms
template <typename T, int N> static constexpr int zfull_occ() {
    if constexpr (is64<T>()) return (N == 32 || N == 64 || N == 256) ? 3 : 2;
    else return (N == 48 || N == 96) ? 2 : 3;
}

// P_z of two lines per forward transform (passz_full_kernel's ZPAIR): measured at 512^3 FP32
// 0.742 -> 0.724 ms per launch; FP64 1.333 -> 1.305 ms alone but the same 1.39 ms inside the step,
// with 10 % more HBM reads (the paired FP64 build spills 36 bytes per lane); at 256^3 FP32
// 0.170 -> 0.178 ms (the paired build spills there).  So only FP32 at 512^3 pairs.
template <typename T, int N> static constexpr bool zfull_pair() { return N == 512 && !is64<T>(); }

template <typename T, int N, int S>
static int launch_passz_rhs(nsg_ctx* c, int b0, int nb, double scale, int track_max) {
    constexpr int EF = zfull_ept<T, N>();
    constexpr int EPTZ = EZof<T, N, S>(), TPFZ = EF ? N / EF : N / 2 / EPTZ, LPW = 64 / TPFZ, WPB = 4;
    static_assert(EPTZ >= 1 && LPW >= 1, "bad z configuration");
    static_assert(Ext<N>::NZP == nzp_of(N), "the kernels' compile-time extents must match the context's");
    PZArgs a;
    a.in = (const char*)c->B2 + b0 * c->a2box * esz<T>();
    a.out = (char*)c->B1 + b0 * c->a3box * esz<T>();
    a.ibox = c->a2box; a.obox = c->a3box; a.nbox = nb;
    a.twM = twM_of<T>(c); a.twN = twN_of<T>(c); a.sc = c->sc + b0;
    a.scale = scale; a.kmax = c->kmax; a.NZP = c->NZP; a.track_max = track_max;
    const long long groups = ((long long)N * N * nb + LPW - 1) / LPW;
    // the full form takes two line groups per wavefront iteration when they pair up
    long long iters = groups;
    if constexpr (EF > 0) {
        if constexpr (zfull_pair<T, N>() && zfull_pairs<N, EF>()) iters = groups / 2;
    }
    long long blocks = (iters + WPB - 1) / WPB;
    const int nf = nsched(N), zg = zgrid_of<T, N>((c->sched / nf) % kNZGrid);
    const long long cap = (long long)c->ncu * (c->zblocks_env > 0 ? c->zblocks_env : zg);
    if (blocks > cap) blocks = cap;
    prof_begin(c);
    if constexpr (EF > 0) {
        hipLaunchKernelGGL((passz_full_kernel<T, N, EF, WPB, zfull_occ<T, N>(), true, zfull_pair<T, N>()>), dim3((unsigned)blocks),
                           dim3(64 * WPB), 0, c->s, a);
    } else {
        // Fields of at most 16 VGPRs per thread: compiled for 2 waves per SIMD.  The kernel's natural
        // allocation sits at 250-256 VGPRs, and a build one register over would fall to 1 wave per
        // SIMD (measured 1.7x slower at 512^3 FP32).  Larger fields (FP64, 8 per thread) would spill
        // under that bound and keep the unbounded build.
        constexpr int OCC = (EPTZ * sizeof(cpx_t<T>) <= 64) ? 2 : 1;
        hipLaunchKernelGGL((passz_rhs_kernel<T, N, EPTZ, WPB, OCC>), dim3((unsigned)blocks), dim3(64 * WPB), 0, c->s, a);
    }
    LAUNCHCHK(c);
    prof_end(c, 2);
    return NSG_OK;
}

// block-synchronous z pass (every mode); in/out point at the first box, ibox/obox are per-box
// strides in elements of the input/output type
template <typename T, int N, int S, int MODE>
static int launch_passz(nsg_ctx* c, const void* in, void* out, long long ibox, long long obox, int nb, DevScalars* sc,
                        double scale, int track_max, int soa = 0) {
    constexpr int EPTZ = EZof<T, N, S>(), TPFZ = N / 2 / EPTZ;
    constexpr int LPB = (256 / TPFZ) > 0 ? (256 / TPFZ) : 1;
    PZArgs a;
    a.in = in; a.out = out; a.ibox = ibox; a.obox = obox; a.nbox = nb;
    a.twM = twM_of<T>(c); a.twN = twN_of<T>(c); a.sc = sc;
    a.scale = scale; a.kmax = c->kmax; a.NZP = c->NZP; a.track_max = track_max; a.soa = soa;
    const long long lines = (long long)N * N * nb;
    const unsigned blocks = (unsigned)((lines + LPB - 1) / LPB);
    prof_begin(c);
    hipLaunchKernelGGL((passz_kernel<T, N, EPTZ, LPB, MODE>), dim3(blocks), dim3(TPFZ * LPB), 0, c->s, a);
    LAUNCHCHK(c);
    prof_end(c, 2);
    return NSG_OK;
}

// DPF: a Dormand-Prince family context (pass5_kernel FAM 1 in its own layout)
template <typename T, int N, int S, int MODE, bool DPF = false>
static int launch_pass5(nsg_ctx* c, double2* uh, int b0, int nb, int stage, void* kout, int force, double norm,
                        double2* unew = nullptr, double* errpart = nullptr) {
    constexpr int EPT = DPF ? E5dp<T, N, S>() : E5of<T, N, S>(), W = DPF ? W5dp<T, N, S>() : W5of<T, N, S>();
    constexpr int TPF = N / EPT;
    static_assert(TPF * W <= 1024, "block too large");
    constexpr size_t e = esz<T>();
    P5Args a;
    a.in = (const char*)c->B2 + b0 * c->a4box * e;
    a.uh = uh + b0 * c->sbox;
    a.unew = unew ? unew + b0 * c->sbox : nullptr;
    fill_row<T>(c, a.row, b0, stage);
    a.kout = kout ? (char*)kout + b0 * c->sbox * e : nullptr;
    a.nslots = c->NK * (c->NZP / W);
    a.errpart = errpart ? errpart + (long long)b0 * a.nslots * 3 : nullptr;
    a.nstage = c->nstage;
    a.extra = (MODE == EP_STAGE && c->closure && c->cnn) ? c->closure + b0 * c->sbox : nullptr;
    a.twN = twN_of<T>(c); a.sc = c->sc + b0; a.fslots = c->fslots + (long long)b0 * c->NK;
    a.kf2 = c->p.force_kf * c->p.force_kf; a.eps_in = c->p.force_eps; a.norm = norm;
    a.sbox = c->sbox; a.ibox = c->a4box;
    a.visc_pow = c->p.visc_pow; a.scheme = c->p.scheme; a.stage = stage; a.force = force;
    a.kmax = c->kmax; a.NK = c->NK; a.NZP = c->NZP;
    const dim3 grid = xgrid<T, W>(c->NZP / W, c->NK, nb);
    prof_begin(c);
    if constexpr (MODE == EP_STAGE && !DPF && p5_pipelined<T, N, W>()) {
        {
            const int ntiles = c->NZP / W;
            const long long nitems = (long long)ntiles * c->NK * nb;
            const int blocks = (int)std::min<long long>(nitems, c->ncu);  // 2 N W elements of LDS: one block per CU
            if (c->p.scheme == RK4CN)
                hipLaunchKernelGGL((pass5p_kernel<T, N, EPT, W, RK4CN>), dim3((unsigned)blocks), dim3(TPF * W), 0, c->s, a,
                                   ntiles, (int)nitems);
            else
                hipLaunchKernelGGL((pass5p_kernel<T, N, EPT, W, IFRK4>), dim3((unsigned)blocks), dim3(TPF * W), 0, c->s, a,
                                   ntiles, (int)nitems);
            LAUNCHCHK(c);
            prof_end(c, 4);
            return NSG_OK;
        }
    }
    if constexpr (MODE == EP_STAGE) {
        if constexpr (DPF)
            hipLaunchKernelGGL((pass5_kernel<T, N, EPT, W, MODE, 1>), grid, dim3(TPF * W), 0, c->s, a);
        else
            hipLaunchKernelGGL((pass5_kernel<T, N, EPT, W, MODE, 0>), grid, dim3(TPF * W), 0, c->s, a);
    } else {
        hipLaunchKernelGGL((pass5_kernel<T, N, EPT, W, MODE, 0>), grid, dim3(TPF * W), 0, c->s, a);
    }
    LAUNCHCHK(c);
    prof_end(c, 4);
    return NSG_OK;
}

// one RK stage s (1..4) of the nonlinear term on boxes [b0, b0+nb) of `uh`, writing K_s (s<4)
// or the update (final_update)
template <typename T, int N, int S>
static int run_stage(nsg_ctx* c, double2* uh, int b0, int nb, int stage, bool track_max, bool final_update) {
    const int force = c->p.force_eps > 0.0 ? 1 : 0;
    const void* kprev = stage > 1 ? c->K[stage - 2] : nullptr;
    const double N3 = (double)N * N * N;
    int rc;
    if ((rc = launch_pass1<T, N, S>(c, uh, b0, nb, kprev, stage, 0, force))) return rc;
    if ((rc = launch_passy<T, N, S, +1>(c, b0, nb, 6))) return rc;
    if constexpr (zfull_ept<T, N>() > 0 || zrhs_persistent<T, N, S>()) {
        if ((rc = launch_passz_rhs<T, N, S>(c, b0, nb, 1.0 / N3, track_max ? 1 : 0))) return rc;
    } else {
        if ((rc = launch_passz<T, N, S, ZM_RHS>(c, (const char*)c->B2 + b0 * c->a2box * esz<T>(),
                                               (char*)c->B1 + b0 * c->a3box * esz<T>(), c->a2box, c->a3box, nb,
                                               c->sc + b0, 1.0 / N3, track_max ? 1 : 0)))
            return rc;
    }
    if (track_max && c->any_cfl) {
        hipLaunchKernelGGL(cfl_kernel, dim3(1), dim3(64), 0, c->s, c->sc + b0, nb);
        LAUNCHCHK(c);
    }
    if ((rc = launch_passy<T, N, S, -1>(c, b0, nb, 3))) return rc;
    void* kout = final_update ? nullptr : c->K[stage - 1];
    const bool dp = final_update && c->p.scheme == NSG_SCHEME_DP5;
    if (dp_family(c))
        return launch_pass5<T, N, S, EP_STAGE, true>(c, uh, b0, nb, final_update ? c->nstage : stage, kout, force,
                                                     1.0 / (N3 * N3), dp ? c->ucand : nullptr, dp ? c->dperr : nullptr);
    return launch_pass5<T, N, S, EP_STAGE>(c, uh, b0, nb, final_update ? c->nstage : stage, kout, force,
                                           1.0 / (N3 * N3), nullptr, nullptr);
}

// one step of boxes [b0, b0+nb); Dormand-Prince: one try, whose candidate lands in ucand and
// whose error norm and step control land in the box scalars (the caller accepts by swapping)
template <typename T, int N, int S> static int run_step(nsg_ctx* c, int b0, int nb) {
    int rc;
    hipLaunchKernelGGL(step_begin_kernel, dim3(1), dim3(64), 0, c->s, c->sc + b0, nb);
    LAUNCHCHK(c);
    for (int s = 1; s <= c->nstage; ++s)
        if ((rc = run_stage<T, N, S>(c, c->uh, b0, nb, s, s == 1, s == c->nstage))) return rc;
    if (c->p.scheme == NSG_SCHEME_DP5) {
        constexpr int W = W5dp<T, N, S>();
        const int nslots = c->NK * (c->NZP / W);
        const double count = (double)N * N * (N / 2 + 1);
        hipLaunchKernelGGL(dp_finalize_kernel, dim3(nb), dim3(256), 0, c->s, c->sc + b0,
                           c->dperr + (long long)b0 * nslots * 3, nslots, count);
    } else {
        hipLaunchKernelGGL(step_end_kernel, dim3(1), dim3(64), 0, c->s, c->sc + b0, nb);
    }
    LAUNCHCHK(c);
    return NSG_OK;
}

// enqueue n steps of boxes [b0, b0+nb) on the batch's stream
static int enqueue_steps(nsg_ctx* c, long n, int b0, int nb) {
    return dispatch_NS(c->N, fft_sched(c), [&](auto nn, auto ss) {
        constexpr int NN = decltype(nn)::value, SS = decltype(ss)::value;
        for (long i = 0; i < n; ++i) {
            const int rc = c->fp32 ? run_step<float, NN, SS>(c, b0, nb) : run_step<double, NN, SS>(c, b0, nb);
            if (rc) return rc;
        }
        return (int)NSG_OK;
    });
}

// after steps: wait for the stream, then check every box's scalars (finite, under its limit)
static int check_boxes(nsg_ctx* c) {
    std::vector<DevScalars> hs(c->nbox);
    HIPCHK(c, hipMemcpyAsync(hs.data(), c->sc, sizeof(DevScalars) * c->nbox, hipMemcpyDeviceToHost, c->s));
    HIPCHK(c, hipStreamSynchronize(c->s));
    for (int b = 0; b < c->nbox; ++b) {
        const char* box = c->nbox > 1 ? " in box " : "";
        char bs[16] = "";
        if (c->nbox > 1) snprintf(bs, sizeof bs, "%d", b);
        if (!std::isfinite(hs[b].max_last) || !std::isfinite(hs[b].dt))
            return fail(c, NSG_ERR_NONFINITE, "non-finite velocity%s%s after step %lld", box, bs, hs[b].steps);
        if (c->max_speed[b] > 0.0 && hs[b].max_last > c->max_speed[b])
            return fail(c, NSG_ERR_BLOWUP, "max|u| = %g exceeds the limit %g%s%s at step %lld", hs[b].max_last,
                        c->max_speed[b], box, bs, hs[b].steps);
    }
    return NSG_OK;
}

// ------------------------------------------------------------------------------------------
// batch creation
// ------------------------------------------------------------------------------------------
static void free_all(nsg_ctx* c) {
    if (!c) return;
    nsgi_closure_free(c);
    void* bufs[] = {c->uh, c->aux, c->K[0], c->K[1], c->K[2], c->K[3], c->K[4], c->K[5], c->ucand, c->dperr,
                    c->B1, c->B2, c->twN64, c->twM64, c->twN32, c->twM32, c->sc, c->fslots, c->part};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    if (c->ev[0]) (void)hipEventDestroy(c->ev[0]);
    if (c->ev[1]) (void)hipEventDestroy(c->ev[1]);
    if (c->s) (void)hipStreamDestroy(c->s);
}

static const char* check_params(const nsg_params* p) {
    const int N = p->N[0];
    if (p->N[1] != N || p->N[2] != N) return "only cubic grids are supported (N[0]==N[1]==N[2])";
    if (dispatch_N(N, [](auto) { return 0; }) != 0) return "unsupported N (8,12,16,24,32,48,64,96,128,256,512,1024)";
    if (!(p->nu >= 0.0) || !(p->dt > 0.0)) return "nu must be >= 0 and dt > 0";
    if (p->visc_pow < 1 || p->visc_pow > 4) return "visc_pow must be 1..4";
    if (p->scheme < NSG_SCHEME_RK4CN || p->scheme > NSG_SCHEME_DP5) return "unknown scheme";
    if (p->scheme == NSG_SCHEME_DP5 && p->cfl > 0.0) return "the Dormand-Prince scheme chooses dt itself: cfl must be 0";
    // FP32 round-off of the slopes at low-amplitude modes exceeds the 1e-7 per-mode tolerance
    // (measured: the error norm of one try reads 8% high at 32^3, 2.3x at 128^3)
    if (p->scheme == NSG_SCHEME_DP5 && p->precision != NSG_FP64)
        return "the Dormand-Prince error norm (tolerance 1e-7 per mode) is below FP32 round-off: use FP64";
    if (p->precision != NSG_FP64 && p->precision != NSG_FP32BULK) return "unknown precision";
    if (p->force_eps > 0.0 && !(p->force_kf > 0.0 && p->force_kf < 7.5)) return "force_kf must be in (0, 7.5)";
    return nullptr;
}

// the parameters every box of one batch must share
static bool same_batch(const nsg_params& a, const nsg_params& b) {
    return a.N[0] == b.N[0] && a.N[1] == b.N[1] && a.N[2] == b.N[2] && a.precision == b.precision &&
           a.scheme == b.scheme && a.visc_pow == b.visc_pow && a.force_eps == b.force_eps &&
           a.force_kf == b.force_kf && a.device == b.device;
}

// boxes p[0..nbox) share N, precision, scheme, visc_pow, forcing and device; nu, dt, t0, cfl and
// max_speed may differ per box
static int create_batch(const nsg_params* p, int nbox, nsg_ctx** out) {
    if (!p || !out || nbox < 1) return NSG_ERR_INVALID;
    *out = nullptr;
    nsg_ctx* c = new (std::nothrow) nsg_ctx();
    if (!c) return NSG_ERR_OOM;
    *out = c;  // handed back on failure too, so the caller can read nsg_last_error, then destroy it
    c->p = p[0];
    for (int b = 0; b < nbox; ++b) {
        if (const char* m = check_params(&p[b])) return fail(c, NSG_ERR_INVALID, "%s", m);
        if (!same_batch(p[b], p[0]))
            return fail(c, NSG_ERR_INVALID, "boxes of one batch must share N, precision, scheme, visc_pow, forcing, device");
    }
    if (p[0].scheme == NSG_SCHEME_DP5 && nbox != 1)
        return fail(c, NSG_ERR_INVALID, "the adaptive Dormand-Prince scheme runs one box per context");
    const int N = p[0].N[0];
    c->N = N;
    c->nbox = nbox;
    c->nslopes = p[0].scheme == NSG_SCHEME_DP5 ? 6 : p[0].scheme == NSG_SCHEME_RK5CN ? 5 : 3;
    c->nstage = p[0].scheme == NSG_SCHEME_DP5 ? 7 : p[0].scheme == NSG_SCHEME_RK5CN ? 6 : 4;
    c->kmax = N / 3;
    c->NK = 2 * c->kmax + 1;
    c->NZP = nzp_of(N);
    c->nspect = (int)std::sqrt(3.0 * (N / 2.0) * (N / 2.0)) + 1;
    c->fp32 = p[0].precision == NSG_FP32BULK;
    c->max_speed.resize(nbox);
    for (int b = 0; b < nbox; ++b) {
        c->max_speed[b] = p[b].max_speed;
        if (p[b].cfl > 0.0) c->any_cfl = true;
    }

    hipError_t e = hipSetDevice(p[0].device);
    if (e != hipSuccess) return fail(c, NSG_ERR_DEVICE, "hipSetDevice(%d): %s", p[0].device, hipGetErrorString(e));
    HIPCHK(c, hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking));
    nsgi_noalloc_setup(p[0].device);
    {
        hipDeviceProp_t prop;
        HIPCHK(c, hipGetDeviceProperties(&prop, p[0].device));
        c->ncu = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
        const char* zb = std::getenv("NSG_ZBLOCKS_PER_CU");
        if (zb && std::atoi(zb) > 0) c->zblocks_env = std::atoi(zb);
    }
    HIPCHK(c, hipEventCreate(&c->ev[0]));
    HIPCHK(c, hipEventCreate(&c->ev[1]));

    const size_t NK = c->NK, NZP = c->NZP, Nl = N, B = nbox;
    c->sbox = (long long)(3 * NK * NK * NZP);
    c->a1box = 6LL * N * xplane(c->NK, c->NZP);
    c->a2box = 6LL * N * N * c->NZP;
    c->a3box = 3LL * N * N * c->NZP;
    c->a4box = 3LL * N * xplane(c->NK, c->NZP);
    const size_t nstate = (size_t)c->sbox * B;
    const size_t ebulk = c->fp32 ? sizeof(float2) : sizeof(double2);
    // B1: A1 / A3 of every box (sized for FP64: the IC and real-space paths run in FP64), or one
    //     box's real AoS field [N^3][3]
    size_t b1 = (size_t)std::max(c->a1box, c->a3box) * B * sizeof(double2);
    b1 = std::max(b1, 3 * Nl * Nl * Nl * sizeof(double));
    // B2: A2 / A4 of every box, or the AoS spectral staging [N][N][N/2+1][3], or a real AoS field
    size_t b2 = (size_t)std::max(c->a2box, c->a4box) * B * sizeof(double2);
    b2 = std::max(b2, 3 * Nl * Nl * (Nl / 2 + 1) * sizeof(double2));
    b2 = std::max(b2, 3 * Nl * Nl * Nl * sizeof(double));
    auto alloc = [&](void** ptr, size_t bytes) -> int {
        hipError_t ee = hipMalloc(ptr, bytes);
        if (ee != hipSuccess) return fail(c, NSG_ERR_OOM, "hipMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(ee));
        return NSG_OK;
    };
    int rc;
    if ((rc = alloc((void**)&c->uh, nstate * sizeof(double2)))) return rc;
    if ((rc = alloc((void**)&c->aux, nstate * sizeof(double2)))) return rc;
    for (int i = 0; i < c->nslopes; ++i)
        if ((rc = alloc(&c->K[i], nstate * ebulk))) return rc;
    if (p[0].scheme == NSG_SCHEME_DP5) {
        if ((rc = alloc((void**)&c->ucand, nstate * sizeof(double2)))) return rc;
        // one slot per final-pass block; W >= 1 bounds every schedule's count
        if ((rc = alloc((void**)&c->dperr, sizeof(double) * 3 * NK * NZP * B))) return rc;
    }
    if ((rc = alloc(&c->B1, b1))) return rc;
    if ((rc = alloc(&c->B2, b2))) return rc;
    if ((rc = alloc((void**)&c->sc, sizeof(DevScalars) * B))) return rc;
    if ((rc = alloc((void**)&c->fslots, sizeof(double) * NK * B))) return rc;
    c->part_n = std::max<size_t>(7 * (size_t)c->nspect, ((size_t)c->sbox / 3 + 255) / 256 + 1);
    if ((rc = alloc((void**)&c->part, sizeof(double) * c->part_n))) return rc;
    if ((rc = upload_table(c, N, &c->twN64, &c->twN32))) return rc;
    if ((rc = upload_table(c, N / 2, &c->twM64, &c->twM32))) return rc;
    HIPCHK(c, hipMemset(c->uh, 0, nstate * sizeof(double2)));
    HIPCHK(c, hipMemset(c->fslots, 0, sizeof(double) * NK * B));
    for (int i = 0; i < c->nslopes; ++i) HIPCHK(c, hipMemset(c->K[i], 0, nstate * ebulk));
    if (c->ucand) HIPCHK(c, hipMemset(c->ucand, 0, nstate * sizeof(double2)));
    std::vector<DevScalars> hs(B);
    for (int b = 0; b < nbox; ++b) {
        hs[b] = DevScalars{};
        hs[b].dt = p[b].dt; hs[b].t = p[b].t0; hs[b].maxbits = 0; hs[b].max_last = 0.0; hs[b].steps = 0;
        hs[b].cfl = p[b].cfl; hs[b].dx = 2.0 * M_PI / N; hs[b].nu = p[b].nu;
    }
    HIPCHK(c, hipMemcpy(c->sc, hs.data(), sizeof(DevScalars) * B, hipMemcpyHostToDevice));
    HIPCHK(c, hipDeviceSynchronize());  // the null-stream memsets above must land before work on c->s
    return NSG_OK;
}

// ------------------------------------------------------------------------------------------
// per-box operations (box b of a batch); each synchronises the batch's stream before returning
// ------------------------------------------------------------------------------------------
static size_t aos_spectral_bytes(const nsg_ctx* c) { return 3 * (size_t)c->N * c->N * (c->N / 2 + 1) * sizeof(double2); }
static size_t aos_real_bytes(const nsg_ctx* c) { return 3 * (size_t)c->N * c->N * c->N * sizeof(double); }

// host u_hat [N][N][N/2+1][3] -> state box b of `dst` (masked to the 2/3 sphere)
static int set_state_box(nsg_ctx* c, int b, const double* u_hat, double2* dst) {
    HIPCHK(c, hipMemcpyAsync(c->B2, u_hat, aos_spectral_bytes(c), hipMemcpyHostToDevice, c->s));
    hipLaunchKernelGGL(aos_to_state_kernel, dim3(2048), dim3(256), 0, c->s, (const double2*)c->B2, dst + b * c->sbox,
                       c->N, c->kmax, c->NK, c->NZP);
    LAUNCHCHK(c);
    HIPCHK(c, hipStreamSynchronize(c->s));
    return NSG_OK;
}

// one box's compact array (state: FP64; K: bulk precision) -> host AoS [N][N][N/2+1][3]
static int download_compact(nsg_ctx* c, const void* src, bool bulk, double* out) {
    if (bulk && c->fp32)
        hipLaunchKernelGGL(compact_to_aos_kernel<float2>, dim3(2048), dim3(256), 0, c->s, (const float2*)src, (double2*)c->B2, c->N, c->kmax, c->NK, c->NZP);
    else
        hipLaunchKernelGGL(compact_to_aos_kernel<double2>, dim3(2048), dim3(256), 0, c->s, (const double2*)src, (double2*)c->B2, c->N, c->kmax, c->NK, c->NZP);
    LAUNCHCHK(c);
    HIPCHK(c, hipMemcpyAsync(out, c->B2, aos_spectral_bytes(c), hipMemcpyDeviceToHost, c->s));
    HIPCHK(c, hipStreamSynchronize(c->s));
    return NSG_OK;
}

// r2c of a physical real AoS field [N^3][3] at src into the compact FP64 array dst of box b
// (FP64, schedule 0).  A3/A4 of box b are the scratch: src must not lie in B1; it may be the
// start of B2, which pass 3 has consumed before pass 4 writes A4 there.
template <int N> static int r2c_to(nsg_ctx* c, const double* src, double2* dst, int b) {
    int rc;
    if ((rc = launch_passz<double, N, 0, ZM_R2C>(c, src, (double2*)c->B1 + b * c->a3box, 0, 0, 1, c->sc + b, 1.0, 0)))
        return rc;
    if ((rc = launch_passy<double, N, 0, -1>(c, b, 1, 3))) return rc;
    return launch_pass5<double, N, 0, EP_STORE>(c, dst, b, 1, 1, nullptr, 0, 1.0);
}
template <int N> static int r2c_into_state(nsg_ctx* c, int b) { return r2c_to<N>(c, (const double*)c->B2, c->uh, b); }

// c2r of box b's compact FP64 3-vector (pointer to box 0 of its array) to a physical real AoS
// field at dst (FP64, schedule 0); A1 (B1) and A2 (B2) of box b are the scratch
template <int N> static int c2r_from(nsg_ctx* c, const double2* src0, int b, double* dst) {
    int rc;
    if ((rc = launch_pass1<double, N, 0>(c, src0, b, 1, nullptr, 1, 1, 0))) return rc;
    if ((rc = launch_passy<double, N, 0, +1>(c, b, 1, 3))) return rc;
    const double N3 = (double)N * N * N;
    return launch_passz<double, N, 0, ZM_REAL>(c, (const double2*)c->B2 + b * c->a2box, dst, 0, 0, 1, c->sc + b, 1.0 / N3, 0);
}

int nsgi_c2r(nsg_ctx* c, const double2* src, double* dst) {
    return dispatch_N(c->N, [&](auto n) { return c2r_from<decltype(n)::value>(c, src, 0, dst); });
}
int nsgi_r2c(nsg_ctx* c, const double* src, double2* dst) {
    return dispatch_N(c->N, [&](auto n) { return r2c_to<decltype(n)::value>(c, src, dst, 0); });
}

static int init_state_box(nsg_ctx* c, int b, int kind, double k0, double energy, uint64_t seed) {
    if (kind == NSG_IC_TAYLOR_GREEN || kind == NSG_IC_SHAPIRO) {
        hipLaunchKernelGGL(ic_real_kernel, dim3(4096), dim3(256), 0, c->s, (double*)c->B2, c->N,
                           kind == NSG_IC_TAYLOR_GREEN ? 0 : 1, c->p.nu, 0.0);
        LAUNCHCHK(c);
        const int rc = dispatch_N(c->N, [&](auto n) { return r2c_into_state<decltype(n)::value>(c, b); });
        if (rc) return rc;
    } else if (kind == NSG_IC_K4) {
        if (!(k0 > 0.0) || !(energy >= 0.0)) return fail(c, NSG_ERR_INVALID, "k4 IC needs k0 > 0 and energy >= 0");
        const long long plane = (long long)c->NK * c->NK * c->NZP;
        const unsigned nb = (unsigned)((plane + 255) / 256);
        if ((size_t)nb + 1 > c->part_n) return fail(c, NSG_ERR_INVALID, "internal: partial buffer too small");
        double2* uh = c->uh + b * c->sbox;
        hipLaunchKernelGGL(ic_k4_kernel, dim3(nb), dim3(256), 0, c->s, uh, c->part, c->N, c->kmax, c->NK, c->NZP, k0,
                           (unsigned long long)seed);
        LAUNCHCHK(c);
        hipLaunchKernelGGL(sum_parts_kernel, dim3(1), dim3(256), 0, c->s, c->part, (int)nb, c->part + nb);
        LAUNCHCHK(c);
        const double N3 = (double)c->N * c->N * c->N;
        hipLaunchKernelGGL(scale_state_kernel, dim3(2048), dim3(256), 0, c->s, uh, 3 * plane, (const double*)(c->part + nb), energy, N3);
        LAUNCHCHK(c);
    } else {
        return fail(c, NSG_ERR_INVALID, "unknown initial condition %d", kind);
    }
    HIPCHK(c, hipStreamSynchronize(c->s));
    return NSG_OK;
}

static int measures_box(nsg_ctx* c, int b, nsg_totals* out, double* E_k, double* Ens_k, int n_spect, int quirk) {
    if ((E_k || Ens_k) && n_spect < c->nspect) return fail(c, NSG_ERR_INVALID, "n_spect %d < %d", n_spect, c->nspect);
    const int ns = c->nspect;
    hipLaunchKernelGGL(measures_kernel, dim3(ns, 1), dim3(1024), 0, c->s, c->uh + b * c->sbox, c->part, c->N, c->kmax,
                       c->NK, c->NZP, (const DevScalars*)(c->sc + b), c->p.visc_pow, quirk ? 1 : 0, c->sbox);
    LAUNCHCHK(c);
    std::vector<double> h((size_t)ns * 7);
    HIPCHK(c, hipMemcpyAsync(h.data(), c->part, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->s));
    HIPCHK(c, hipStreamSynchronize(c->s));
    const double N3 = (double)c->N * c->N * c->N;
    const double nf = 0.5 / (N3 * N3), cf = 8.0 * M_PI * M_PI * M_PI;
    double t[5] = {0, 0, 0, 0, 0};
    for (int q = 0; q < ns; ++q)
        for (int j = 0; j < 5; ++j) t[j] += h[(size_t)q * 7 + j];
    out->energy = t[0] * cf * nf;
    out->enstrophy = t[1] * cf * nf;
    out->palinstrophy = t[2] * cf * nf;
    out->helicity = t[3] * cf * nf;
    out->dissipation = t[4] * 2.0 * cf * nf;
    for (int q = 0; q < n_spect && (E_k || Ens_k); ++q) {
        if (E_k) E_k[q] = q < ns ? cf * nf * h[(size_t)q * 7 + 5] : 0.0;
        if (Ens_k) Ens_k[q] = q < ns ? cf * nf * h[(size_t)q * 7 + 6] : 0.0;
    }
    const double* vals[] = {&out->energy, &out->enstrophy, &out->palinstrophy, &out->helicity, &out->dissipation};
    for (const double* v : vals)
        if (!std::isfinite(*v)) return fail(c, NSG_ERR_NONFINITE, "non-finite diagnostics");
    return NSG_OK;
}

// real-space u (which 0) or w (which 1) of box b as AoS [N^3][3] doubles (soa: three planes
// [3][N^3]) at the start of B1
template <int N> static int realspace_N(nsg_ctx* c, int b, int which, int soa = 0) {
    int rc;
    if ((rc = launch_pass1<double, N, 0>(c, c->uh, b, 1, nullptr, 1, which == 0 ? 1 : 2, 0))) return rc;
    // pass 2 on three fields: A1 [3][N][NK][NZP] -> A2 [3][N][N][NZP]
    if ((rc = launch_passy<double, N, 0, +1>(c, b, 1, 3))) return rc;
    const double N3 = (double)N * N * N;
    return launch_passz<double, N, 0, ZM_REAL>(c, (const double2*)c->B2 + b * c->a2box, c->B1, 0, 0, 1, c->sc + b,
                                               1.0 / N3, 0, soa);
}

int nsgi_realspace(nsg_ctx* c, const double** field, int soa) {
    HIPCHK(c, hipSetDevice(c->p.device));
    const int rc = dispatch_N(c->N, [&](auto n) { return realspace_N<decltype(n)::value>(c, 0, 0, soa); });
    if (rc) return rc;
    *field = (const double*)c->B1;
    return NSG_OK;
}

static int realspace_box(nsg_ctx* c, int b, int which, double* out) {
    const int rc = dispatch_N(c->N, [&](auto n) { return realspace_N<decltype(n)::value>(c, b, which); });
    if (rc) return rc;
    HIPCHK(c, hipMemcpyAsync(out, c->B1, aos_real_bytes(c), hipMemcpyDeviceToHost, c->s));
    HIPCHK(c, hipStreamSynchronize(c->s));
    return NSG_OK;
}

static int get_time_box(nsg_ctx* c, int b, double* t, double* dt, double* max_speed) {
    DevScalars hs;
    HIPCHK(c, hipMemcpyAsync(&hs, c->sc + b, sizeof hs, hipMemcpyDeviceToHost, c->s));
    HIPCHK(c, hipStreamSynchronize(c->s));
    if (t) *t = hs.t;
    if (dt) *dt = hs.dt;
    if (max_speed) *max_speed = hs.max_last;
    return NSG_OK;
}

static int set_time_box(nsg_ctx* c, int b, double t, double dt) {
    DevScalars hs;
    HIPCHK(c, hipMemcpyAsync(&hs, c->sc + b, sizeof hs, hipMemcpyDeviceToHost, c->s));
    HIPCHK(c, hipStreamSynchronize(c->s));
    hs.t = t;
    hs.dt = dt;
    HIPCHK(c, hipMemcpyAsync(c->sc + b, &hs, sizeof hs, hipMemcpyHostToDevice, c->s));
    HIPCHK(c, hipStreamSynchronize(c->s));
    return NSG_OK;
}

// ------------------------------------------------------------------------------------------
// C ABI, single box (C linkage comes from the declarations in nsg.h)
// ------------------------------------------------------------------------------------------

void nsg_default_params(nsg_params* p) {
    std::memset(p, 0, sizeof *p);
    p->N[0] = p->N[1] = p->N[2] = 64;
    p->nu = 1.0;
    p->visc_pow = 1;
    p->dt = 1e-3;
    p->t0 = 0.0;
    p->scheme = NSG_SCHEME_RK4CN;
    p->precision = NSG_FP64;
    p->cfl = 0.0;
    p->force_eps = 0.0;
    p->force_kf = 2.5;
    p->max_speed = 0.0;
    p->device = 0;
}

const char* nsg_last_error(const nsg_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int nsg_n_spect(const nsg_ctx* ctx) { return ctx ? ctx->nspect : 0; }

void nsg_destroy(nsg_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->p.device);
    if (ctx->s) (void)hipStreamSynchronize(ctx->s);
    free_all(ctx);
    delete ctx;
}

int nsg_create(const nsg_params* p, nsg_ctx** out) { return create_batch(p, 1, out); }

int nsg_set_state(nsg_ctx* c, const double* u_hat) {
    if (!c || !u_hat) return NSG_ERR_INVALID;
    HIPCHK(c, hipSetDevice(c->p.device));
    return set_state_box(c, 0, u_hat, c->uh);
}

int nsg_get_state(nsg_ctx* c, double* u_hat) {
    if (!c || !u_hat) return NSG_ERR_INVALID;
    HIPCHK(c, hipSetDevice(c->p.device));
    return download_compact(c, c->uh, false, u_hat);
}

int nsg_set_state_real(nsg_ctx* c, const double* u) {
    if (!c || !u) return NSG_ERR_INVALID;
    HIPCHK(c, hipSetDevice(c->p.device));
    HIPCHK(c, hipMemcpyAsync(c->B2, u, aos_real_bytes(c), hipMemcpyHostToDevice, c->s));
    const int rc = dispatch_N(c->N, [&](auto n) { return r2c_into_state<decltype(n)::value>(c, 0); });
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->s));
    return NSG_OK;
}

int nsg_init_state(nsg_ctx* c, int kind, double k0, double energy, uint64_t seed) {
    if (!c) return NSG_ERR_INVALID;
    HIPCHK(c, hipSetDevice(c->p.device));
    return init_state_box(c, 0, kind, k0, energy, seed);
}

int nsg_rhs(nsg_ctx* c, const double* u_hat, double* out) {
    if (!c || !u_hat || !out) return NSG_ERR_INVALID;
    HIPCHK(c, hipSetDevice(c->p.device));
    int rc = set_state_box(c, 0, u_hat, c->aux);
    if (rc) return rc;
    rc = dispatch_NS(c->N, fft_sched(c), [&](auto nn, auto ss) {
        constexpr int NN = decltype(nn)::value, SS = decltype(ss)::value;
        return c->fp32 ? run_stage<float, NN, SS>(c, c->aux, 0, 1, 1, false, false)
                       : run_stage<double, NN, SS>(c, c->aux, 0, 1, 1, false, false);
    });
    if (rc) return rc;
    return download_compact(c, c->K[0], true, out);
}

// Dormand-Prince: tries until one is accepted (at most 50, Solver/src/data_types.h:128), each
// followed by a read of the box scalars; an accepted candidate becomes the state by a swap
static int dp_steps(nsg_ctx* c, long nsteps) {
    constexpr int MAX_TRY = 50;
    int rc;
    for (long i = 0; i < nsteps; ++i) {
        if (c->cnn && (rc = nsgi_closure_eval(c))) return rc;
        int tries = 0;
        for (;;) {
            if ((rc = enqueue_steps(c, 1, 0, 1))) return rc;
            DevScalars h;
            HIPCHK(c, hipMemcpyAsync(&h, c->sc, sizeof h, hipMemcpyDeviceToHost, c->s));
            HIPCHK(c, hipStreamSynchronize(c->s));
            if (!std::isfinite(h.dperr) || !std::isfinite(h.dt))
                return fail(c, NSG_ERR_NONFINITE, "non-finite Dormand-Prince error (%g) or dt (%g) at step %lld", h.dperr,
                            h.dt, h.steps);
            if (h.dpacc) {
                std::swap(c->uh, c->ucand);
                if (c->max_speed[0] > 0.0 && h.max_last > c->max_speed[0])
                    return fail(c, NSG_ERR_BLOWUP, "max|u| = %g exceeds the limit %g at step %lld", h.max_last,
                                c->max_speed[0], h.steps);
                break;
            }
            ++c->dp_rejected;
            if (++tries >= MAX_TRY)
                return fail(c, NSG_ERR_STEP, "Dormand-Prince step rejected %d times (last error %g, dt now %g)", MAX_TRY,
                            h.dperr, h.dt);
        }
    }
    return NSG_OK;
}

int nsg_step(nsg_ctx* c, long nsteps) {
    if (!c || nsteps < 0) return NSG_ERR_INVALID;
    HIPCHK(c, hipSetDevice(c->p.device));
    int rc;
    if (c->p.scheme == NSG_SCHEME_DP5) return dp_steps(c, nsteps);
    if (c->cnn) {
        // learned closure: the model term is evaluated from the state at the start of each step
        // and held fixed through its four stages
        for (long i = 0; i < nsteps; ++i) {
            if ((rc = nsgi_closure_eval(c))) return rc;
            if ((rc = enqueue_steps(c, 1, 0, c->nbox))) return rc;
        }
    } else if ((rc = enqueue_steps(c, nsteps, 0, c->nbox))) {
        return rc;
    }
    if (nsteps == 0) {
        HIPCHK(c, hipStreamSynchronize(c->s));
        return NSG_OK;
    }
    return check_boxes(c);
}

int nsg_dp_stats(nsg_ctx* c, long long* accepted, long long* rejected, double* last_err) {
    if (!c) return NSG_ERR_INVALID;
    if (c->p.scheme != NSG_SCHEME_DP5) return fail(c, NSG_ERR_INVALID, "not a Dormand-Prince context");
    HIPCHK(c, hipSetDevice(c->p.device));
    DevScalars h;
    HIPCHK(c, hipMemcpyAsync(&h, c->sc, sizeof h, hipMemcpyDeviceToHost, c->s));
    HIPCHK(c, hipStreamSynchronize(c->s));
    if (accepted) *accepted = h.steps;
    if (rejected) *rejected = c->dp_rejected;
    if (last_err) *last_err = h.dperr;
    return NSG_OK;
}

int nsg_measures(nsg_ctx* c, nsg_totals* out, double* E_k, double* Ens_k, int n_spect, int quirk) {
    if (!c || !out) return NSG_ERR_INVALID;
    HIPCHK(c, hipSetDevice(c->p.device));
    return measures_box(c, 0, out, E_k, Ens_k, n_spect, quirk);
}

int nsg_realspace(nsg_ctx* c, int which, double* out) {
    if (!c || !out || (which != 0 && which != 1)) return NSG_ERR_INVALID;
    HIPCHK(c, hipSetDevice(c->p.device));
    return realspace_box(c, 0, which, out);
}

int nsg_get_time(nsg_ctx* c, double* t, double* dt, double* max_speed) {
    if (!c) return NSG_ERR_INVALID;
    HIPCHK(c, hipSetDevice(c->p.device));
    return get_time_box(c, 0, t, dt, max_speed);
}

int nsg_set_time(nsg_ctx* c, double t, double dt) {
    if (!c || !(dt > 0.0)) return NSG_ERR_INVALID;
    HIPCHK(c, hipSetDevice(c->p.device));
    return set_time_box(c, 0, t, dt);
}

static void pass_bytes(const nsg_ctx* c, double* b /*[5]: stage-averaged, one box*/) {
    const double N = c->N, NK = c->NK, NZL = c->kmax + 1;
    const double E = c->fp32 ? 8.0 : 16.0, S = 16.0;
    const double st = 3 * NK * NK * NZL;  // complex elements of a 3-vector state
    const double ns = c->nstage;
    // slopes read by the stage inputs, summed over the step: one per stage s > 1 (RK4 schemes), the
    // nonzero entries of the tableau rows (Dormand-Prince)
    double kin = ns - 1;
    if (dp_family(c)) {
        kin = 0;
        for (int s = 2; s <= c->nstage; ++s)
            for (int j = 0; j < s - 1; ++j) kin += kDPA[s - 1][j] != 0.0;
    }
    // slopes read by the final update: K1..K3 (RK4), K1,K3,K4,K5 (+K6 for Dormand-Prince)
    const double kfin = dp_family(c) ? (c->nstage == 7 ? 5 : 4) : 3;
    // pass 1: read state (+ slopes for s > 1), write 6 fields
    b[0] = st * S + kin / ns * st * E + 6 * N * NK * NZL * E;
    b[1] = (6 * N * NK * NZL + 6 * N * N * NZL) * E;
    b[2] = (6 * N * N * NZL + 3 * N * N * NZL) * E;
    b[3] = (3 * N * N * NZL + 3 * N * NK * NZL) * E;
    // pass 5: read A4, write K_s (s < nstage) | read state + slopes, write state (final stage)
    b[4] = 3 * N * NK * NZL * E + (ns - 1) / ns * st * E + (st * S + kfin * st * E + st * S) / ns;
}

double nsg_step_bytes(const nsg_ctx* c) {
    if (!c) return 0.0;
    double b[5];
    pass_bytes(c, b);
    return c->nstage * (b[0] + b[1] + b[2] + b[3] + b[4]) * c->nbox;
}

int nsg_profile_passes(nsg_ctx* c, int iters, double* ms, double* bytes) {
    if (!c || iters <= 0 || !ms) return NSG_ERR_INVALID;
    HIPCHK(c, hipSetDevice(c->p.device));
    for (int k = 0; k < 5; ++k) { c->prof_ms[k] = 0.0; c->prof_cnt[k] = 0; }
    HIPCHK(c, hipStreamSynchronize(c->s));
    // Dormand-Prince tries are not applied here (no accept step): keep the box scalars as they were
    const bool dp = c->p.scheme == NSG_SCHEME_DP5;
    std::vector<DevScalars> hs(dp ? c->nbox : 0);
    if (dp) HIPCHK(c, hipMemcpy(hs.data(), c->sc, sizeof(DevScalars) * c->nbox, hipMemcpyDeviceToHost));
    c->prof = true;
    const int rc = enqueue_steps(c, iters, 0, c->nbox);
    c->prof = false;
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->s));
    if (dp) HIPCHK(c, hipMemcpy(c->sc, hs.data(), sizeof(DevScalars) * c->nbox, hipMemcpyHostToDevice));
    double tot = 0.0;
    for (int k = 0; k < 5; ++k) {
        ms[k] = c->prof_cnt[k] ? c->prof_ms[k] / c->prof_cnt[k] : 0.0;
        tot += c->prof_ms[k];
    }
    ms[5] = tot / iters;
    if (bytes) {
        pass_bytes(c, bytes);
        for (int k = 0; k < 5; ++k) bytes[k] *= c->nbox;
    }
    return NSG_OK;
}

// ------------------------------------------------------------------------------------------
// Schedules and the autotuner: every schedule of the batch's size is timed on the batch's own
// state (saved in aux and restored after each trial, scalars too) and the fastest is kept.
// Choices are cached per (N, precision, nbox) for the life of the process.
// ------------------------------------------------------------------------------------------
int nsg_schedule_count(int N) {
    const int n = dispatch_N(N, [](auto nn) { return nsched(decltype(nn)::value); });
    return n > 0 ? n * kNZGrid : 0;
}

int nsg_set_schedule(nsg_ctx* c, int sched) {
    if (!c) return NSG_ERR_INVALID;
    const int ns = nsg_schedule_count(c->N);
    if (sched < 0 || sched >= ns) return fail(c, NSG_ERR_INVALID, "schedule %d out of range [0,%d) for N=%d", sched, ns, c->N);
    c->sched = sched;
    return NSG_OK;
}

static std::mutex g_tune_mu;
static std::map<std::tuple<int, int, int>, int> g_tune_cache;

static int autotune_ctx(nsg_ctx* c, int iters, int use_cache, int* chosen, double* ms_per_step) {
    const auto key = std::make_tuple(c->N, (int)c->fp32, c->nbox);
    const int ns = nsg_schedule_count(c->N);
    if (use_cache) {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tune_cache.find(key);
        if (it != g_tune_cache.end()) {
            c->sched = it->second;
            if (chosen) *chosen = c->sched;
            return NSG_OK;
        }
    }
    const size_t sbytes = (size_t)c->sbox * c->nbox * sizeof(double2);
    std::vector<DevScalars> hs(c->nbox);
    HIPCHK(c, hipMemcpyAsync(c->aux, c->uh, sbytes, hipMemcpyDeviceToDevice, c->s));
    HIPCHK(c, hipMemcpyAsync(hs.data(), c->sc, sizeof(DevScalars) * c->nbox, hipMemcpyDeviceToHost, c->s));
    HIPCHK(c, hipStreamSynchronize(c->s));
    int best = 0;
    double best_ms = 1e300;
    for (int S = 0; S < ns; ++S) {
        c->sched = S;
        int rc = enqueue_steps(c, 1, 0, c->nbox);  // untimed: first launches load the code objects
        if (rc) return rc;
        HIPCHK(c, hipEventRecord(c->ev[0], c->s));
        if ((rc = enqueue_steps(c, iters, 0, c->nbox))) return rc;
        HIPCHK(c, hipEventRecord(c->ev[1], c->s));
        HIPCHK(c, hipEventSynchronize(c->ev[1]));
        float ms = 0.f;
        HIPCHK(c, hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
        const double per = ms / iters;
        if (ms_per_step) ms_per_step[S] = per;
        if (per < best_ms) { best_ms = per; best = S; }
        HIPCHK(c, hipMemcpyAsync(c->uh, c->aux, sbytes, hipMemcpyDeviceToDevice, c->s));
        HIPCHK(c, hipMemcpyAsync(c->sc, hs.data(), sizeof(DevScalars) * c->nbox, hipMemcpyHostToDevice, c->s));
        HIPCHK(c, hipStreamSynchronize(c->s));
    }
    c->sched = best;
    if (chosen) *chosen = best;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune_cache[key] = best;
    return NSG_OK;
}

int nsg_autotune(nsg_ctx* c, int iters, int use_cache, int* chosen, double* ms_per_step) {
    if (!c || iters <= 0) return NSG_ERR_INVALID;
    HIPCHK(c, hipSetDevice(c->p.device));
    return autotune_ctx(c, iters, use_cache, chosen, ms_per_step);
}

// ------------------------------------------------------------------------------------------
// Ensembles: boxes grouped into batches (one per set of shared parameters, in order of first
// appearance); one step of every batch is captured once into a HIP graph -- the batches fork
// from one stream onto their own streams and join back -- and nsg_ensemble_step replays it.
// ------------------------------------------------------------------------------------------
struct nsg_ensemble {
    std::vector<nsg_ctx*> groups;
    std::vector<std::pair<int, int>> where;  // box -> (group, index in the group)
    hipStream_t master = nullptr;
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> join;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    std::string err;
    int device = 0;
    int gsteps = 1;  // ensemble steps held by the graph (one replay advances every box by gsteps)
};

static int efail(nsg_ensemble* e, int code, const std::string& msg) {
    if (e) e->err = msg;
    return code;
}
static int egroup(nsg_ensemble* e, const nsg_ctx* g, int rc) { return efail(e, rc, g ? g->err : std::string("group error")); }

#define EHIP(e, expr)                                                                                  \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess) return efail((e), NSG_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

int nsg_ensemble_create(const nsg_params* boxes, int nbox, nsg_ensemble** out) {
    if (!boxes || nbox < 1 || !out) return NSG_ERR_INVALID;
    *out = nullptr;
    nsg_ensemble* e = new (std::nothrow) nsg_ensemble();
    if (!e) return NSG_ERR_OOM;
    *out = e;  // handed back on failure too (nsg_ensemble_last_error, then destroy)
    e->device = boxes[0].device;
    std::vector<std::vector<int>> members;
    e->where.resize(nbox);
    for (int b = 0; b < nbox; ++b) {
        if (const char* m = check_params(&boxes[b])) return efail(e, NSG_ERR_INVALID, "box " + std::to_string(b) + ": " + m);
        if (boxes[b].scheme == NSG_SCHEME_DP5)
            return efail(e, NSG_ERR_INVALID, "box " + std::to_string(b) +
                                                 ": the adaptive Dormand-Prince scheme runs one box per context (nsg_create)");
        if (boxes[b].device != e->device) return efail(e, NSG_ERR_INVALID, "all boxes of an ensemble must use one device");
        int gi = -1;
        for (size_t g = 0; g < members.size(); ++g)
            if (same_batch(boxes[members[g][0]], boxes[b])) { gi = (int)g; break; }
        if (gi < 0) { members.emplace_back(); gi = (int)members.size() - 1; }
        e->where[b] = {gi, (int)members[gi].size()};
        members[gi].push_back(b);
    }
    EHIP(e, hipSetDevice(e->device));
    for (auto& m : members) {
        std::vector<nsg_params> ps;
        for (int b : m) ps.push_back(boxes[b]);
        nsg_ctx* g = nullptr;
        const int rc = create_batch(ps.data(), (int)ps.size(), &g);
        if (g) e->groups.push_back(g);
        if (rc) return egroup(e, g, rc);
    }
    EHIP(e, hipStreamCreateWithFlags(&e->master, hipStreamNonBlocking));
    EHIP(e, hipEventCreateWithFlags(&e->fork, hipEventDisableTiming));
    EHIP(e, hipEventCreate(&e->ev[0]));
    EHIP(e, hipEventCreate(&e->ev[1]));
    e->join.assign(e->groups.size(), nullptr);
    for (auto& ev : e->join) EHIP(e, hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    return NSG_OK;
}

static void ens_drop_graph(nsg_ensemble* e) {
    if (e->exec) (void)hipGraphExecDestroy(e->exec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    e->exec = nullptr;
    e->graph = nullptr;
}

// n steps of every batch launched directly: the batches fork from the master stream onto their
// own streams and join back (the same kernels and order per batch as the graph)
static int ens_direct(nsg_ensemble* e, long n) {
    EHIP(e, hipEventRecord(e->fork, e->master));
    for (size_t g = 0; g < e->groups.size(); ++g) {
        nsg_ctx* c = e->groups[g];
        EHIP(e, hipStreamWaitEvent(c->s, e->fork, 0));
        const int rc = enqueue_steps(c, n, 0, c->nbox);
        if (rc) return egroup(e, c, rc);
        EHIP(e, hipEventRecord(e->join[g], c->s));
        EHIP(e, hipStreamWaitEvent(e->master, e->join[g], 0));
    }
    return NSG_OK;
}

static int ens_build_graph(nsg_ensemble* e) {
    if (e->exec) return NSG_OK;
    if (const char* gs = std::getenv("NSG_ENS_GRAPH_STEPS")) e->gsteps = std::max(1, std::atoi(gs));
    EHIP(e, hipStreamBeginCapture(e->master, hipStreamCaptureModeThreadLocal));
    int rc = NSG_OK;
    std::string msg;
    hipError_t st = hipEventRecord(e->fork, e->master);
    if (st != hipSuccess) { rc = NSG_ERR_DEVICE; msg = std::string("hipEventRecord: ") + hipGetErrorString(st); }
    for (size_t g = 0; g < e->groups.size() && rc == NSG_OK; ++g) {
        nsg_ctx* c = e->groups[g];
        if ((st = hipStreamWaitEvent(c->s, e->fork, 0)) != hipSuccess) {
            rc = NSG_ERR_DEVICE; msg = std::string("hipStreamWaitEvent: ") + hipGetErrorString(st); break;
        }
        if ((rc = enqueue_steps(c, e->gsteps, 0, c->nbox))) { msg = c->err; break; }
        if ((st = hipEventRecord(e->join[g], c->s)) != hipSuccess ||
            (st = hipStreamWaitEvent(e->master, e->join[g], 0)) != hipSuccess) {
            rc = NSG_ERR_DEVICE; msg = std::string("join: ") + hipGetErrorString(st); break;
        }
    }
    hipGraph_t graph = nullptr;
    st = hipStreamEndCapture(e->master, &graph);
    if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return efail(e, rc, msg);
    }
    if (st != hipSuccess) return efail(e, NSG_ERR_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(st));
    e->graph = graph;
    EHIP(e, hipGraphInstantiate(&e->exec, e->graph, nullptr, nullptr, 0));
    return NSG_OK;
}

static int ens_check(nsg_ensemble* e) {
    for (nsg_ctx* g : e->groups) {
        const int rc = check_boxes(g);
        if (rc) return egroup(e, g, rc);
    }
    return NSG_OK;
}

int nsg_ensemble_step(nsg_ensemble* e, long nsteps) {
    if (!e || nsteps < 0) return NSG_ERR_INVALID;
    EHIP(e, hipSetDevice(e->device));
    int rc = ens_build_graph(e);
    if (rc) return rc;
    for (long i = 0; i + e->gsteps <= nsteps; i += e->gsteps) EHIP(e, hipGraphLaunch(e->exec, e->master));
    if (nsteps % e->gsteps && (rc = ens_direct(e, nsteps % e->gsteps))) return rc;
    EHIP(e, hipStreamSynchronize(e->master));
    return ens_check(e);
}

int nsg_ensemble_step_serial(nsg_ensemble* e, long nsteps) {
    if (!e || nsteps < 0) return NSG_ERR_INVALID;
    EHIP(e, hipSetDevice(e->device));
    for (long i = 0; i < nsteps; ++i)
        for (nsg_ctx* g : e->groups)
            for (int b = 0; b < g->nbox; ++b) {
                const int rc = enqueue_steps(g, 1, b, 1);
                if (rc) return egroup(e, g, rc);
            }
    return ens_check(e);
}

int nsg_ensemble_autotune(nsg_ensemble* e, int iters, int use_cache, int* chosen_per_group) {
    if (!e || iters <= 0) return NSG_ERR_INVALID;
    EHIP(e, hipSetDevice(e->device));
    ens_drop_graph(e);
    for (size_t g = 0; g < e->groups.size(); ++g) {
        int ch = 0;
        const int rc = autotune_ctx(e->groups[g], iters, use_cache, &ch, nullptr);
        if (rc) return egroup(e, e->groups[g], rc);
        if (chosen_per_group) chosen_per_group[g] = ch;
    }
    return NSG_OK;
}

int nsg_ensemble_set_schedule(nsg_ensemble* e, int group, int sched) {
    if (!e || group < 0 || group >= (int)e->groups.size()) return NSG_ERR_INVALID;
    const int rc = nsg_set_schedule(e->groups[group], sched);
    if (rc) return egroup(e, e->groups[group], rc);
    ens_drop_graph(e);
    return NSG_OK;
}

int nsg_ensemble_group_count(const nsg_ensemble* e) { return e ? (int)e->groups.size() : 0; }

int nsg_ensemble_group_of(const nsg_ensemble* e, int box, int* group, int* index) {
    if (!e || box < 0 || box >= (int)e->where.size()) return NSG_ERR_INVALID;
    if (group) *group = e->where[box].first;
    if (index) *index = e->where[box].second;
    return NSG_OK;
}

double nsg_ensemble_step_bytes(const nsg_ensemble* e) {
    double s = 0.0;
    if (e)
        for (const nsg_ctx* g : e->groups) s += nsg_step_bytes(g);
    return s;
}

#define ENS_BOX(e, box)                                                             \
    if (!(e) || (box) < 0 || (box) >= (int)(e)->where.size()) return NSG_ERR_INVALID; \
    EHIP((e), hipSetDevice((e)->device));                                           \
    nsg_ctx* g = (e)->groups[(e)->where[(box)].first];                              \
    const int gb = (e)->where[(box)].second

int nsg_ensemble_init_state(nsg_ensemble* e, int box, int kind, double k0, double energy, uint64_t seed) {
    ENS_BOX(e, box);
    const int rc = init_state_box(g, gb, kind, k0, energy, seed);
    return rc ? egroup(e, g, rc) : NSG_OK;
}

int nsg_ensemble_set_state(nsg_ensemble* e, int box, const double* u_hat) {
    ENS_BOX(e, box);
    if (!u_hat) return NSG_ERR_INVALID;
    const int rc = set_state_box(g, gb, u_hat, g->uh);
    return rc ? egroup(e, g, rc) : NSG_OK;
}

int nsg_ensemble_get_state(nsg_ensemble* e, int box, double* u_hat) {
    ENS_BOX(e, box);
    if (!u_hat) return NSG_ERR_INVALID;
    const int rc = download_compact(g, g->uh + gb * g->sbox, false, u_hat);
    return rc ? egroup(e, g, rc) : NSG_OK;
}

int nsg_ensemble_measures(nsg_ensemble* e, int box, nsg_totals* out, double* E_k, double* Ens_k, int n_spect, int quirk) {
    ENS_BOX(e, box);
    if (!out) return NSG_ERR_INVALID;
    const int rc = measures_box(g, gb, out, E_k, Ens_k, n_spect, quirk);
    return rc ? egroup(e, g, rc) : NSG_OK;
}

int nsg_ensemble_realspace(nsg_ensemble* e, int box, int which, double* out) {
    ENS_BOX(e, box);
    if (!out || (which != 0 && which != 1)) return NSG_ERR_INVALID;
    const int rc = realspace_box(g, gb, which, out);
    return rc ? egroup(e, g, rc) : NSG_OK;
}

int nsg_ensemble_get_time(nsg_ensemble* e, int box, double* t, double* dt, double* max_speed) {
    ENS_BOX(e, box);
    const int rc = get_time_box(g, gb, t, dt, max_speed);
    return rc ? egroup(e, g, rc) : NSG_OK;
}

int nsg_ensemble_n_spect(const nsg_ensemble* e, int box) {
    if (!e || box < 0 || box >= (int)e->where.size()) return 0;
    return e->groups[e->where[box].first]->nspect;
}

// wall time of nsteps graph replays on the ensemble stream (HIP events), for benchmarks
int nsg_ensemble_time_steps(nsg_ensemble* e, long nsteps, int serial, double* ms) {
    if (!e || nsteps <= 0 || !ms) return NSG_ERR_INVALID;
    EHIP(e, hipSetDevice(e->device));
    int rc;
    if (serial == 0 && (rc = ens_build_graph(e))) return rc;
    for (nsg_ctx* g : e->groups) EHIP(e, hipStreamSynchronize(g->s));
    EHIP(e, hipEventRecord(e->ev[0], e->master));
    if (serial == 2) {
        if ((rc = ens_direct(e, nsteps))) return rc;
    } else if (serial) {
        // the serial path runs on the groups' streams: order them after the start event and
        // bring them back to the master stream before the end event
        for (nsg_ctx* g : e->groups) EHIP(e, hipStreamWaitEvent(g->s, e->ev[0], 0));
        for (long i = 0; i < nsteps; ++i)
            for (nsg_ctx* g : e->groups)
                for (int b = 0; b < g->nbox; ++b)
                    if ((rc = enqueue_steps(g, 1, b, 1))) return egroup(e, g, rc);
        for (size_t gi = 0; gi < e->groups.size(); ++gi) {
            EHIP(e, hipEventRecord(e->join[gi], e->groups[gi]->s));
            EHIP(e, hipStreamWaitEvent(e->master, e->join[gi], 0));
        }
    } else {
        for (long i = 0; i + e->gsteps <= nsteps; i += e->gsteps) EHIP(e, hipGraphLaunch(e->exec, e->master));
        if (nsteps % e->gsteps && (rc = ens_direct(e, nsteps % e->gsteps))) return rc;
    }
    EHIP(e, hipEventRecord(e->ev[1], e->master));
    EHIP(e, hipEventSynchronize(e->ev[1]));
    float f = 0.f;
    EHIP(e, hipEventElapsedTime(&f, e->ev[0], e->ev[1]));
    *ms = f;
    return ens_check(e);
}

const char* nsg_ensemble_last_error(const nsg_ensemble* e) { return e ? e->err.c_str() : "null ensemble"; }

void nsg_ensemble_destroy(nsg_ensemble* e) {
    if (!e) return;
    (void)hipSetDevice(e->device);
    if (e->master) (void)hipStreamSynchronize(e->master);
    ens_drop_graph(e);
    for (nsg_ctx* g : e->groups) nsg_destroy(g);
    for (auto ev : e->join)
        if (ev) (void)hipEventDestroy(ev);
    if (e->fork) (void)hipEventDestroy(e->fork);
    if (e->ev[0]) (void)hipEventDestroy(e->ev[0]);
    if (e->ev[1]) (void)hipEventDestroy(e->ev[1]);
    if (e->master) (void)hipStreamDestroy(e->master);
    delete e;
}
