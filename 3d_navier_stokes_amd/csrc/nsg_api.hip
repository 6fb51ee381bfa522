// nsg_api.hip -- C ABI implementation (include/nsg.h): context, launch sequencing, transfers.
//
// One RK4 step = 4 stages x 5 passes (see nsg_kernels.hpp) + three one-thread bookkeeping
// kernels, all on the context's stream; nothing is copied to the host inside nsg_step except the
// final status scalars.  The reference counterpart is SpectralSolve's loop body
// (Solver/src/solver.c:125-194) calling RK4Step (:505) -> NonlinearRHSBatch (:620).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/nsg.h"
#include "nsg_kernels.hpp"

using namespace nsg;

// ------------------------------------------------------------------------------------------
// per-N launch configuration
//   EPT   : elements per thread of the length-N column transforms (x and y passes)
//   EPTZ64/EPTZ32 : elements per thread of the length-N/2 line transform (z pass)
//   W64/W32 : kz columns per block in the x/y passes (128-byte row segments)
// ------------------------------------------------------------------------------------------
template <int N> struct Cfg;
// EX/EY/EZ: elements per thread of the x, y (length N) and z (length N/2) transforms;
// WX/WY: kz columns per block of the x and y passes (FP64 / FP32)
#define NSG_CFG(n, ex64, ex32, ey64, ey32, ez64, ez32)                       \
    template <> struct Cfg<n> {                                              \
        static constexpr int EX64 = ex64, EX32 = ex32, EY64 = ey64, EY32 = ey32; \
        static constexpr int EZ64 = ez64, EZ32 = ez32;                       \
        static constexpr int WX64 = 8, WX32 = 16, WY64 = 8, WY32 = 16;       \
    };
//      N     EX64 EX32 EY64 EY32 EZ64 EZ32
NSG_CFG(8,    8,   8,   8,   8,   4,   4)
NSG_CFG(12,   12,  12,  12,  12,  6,   6)
NSG_CFG(16,   8,   8,   8,   8,   4,   8)
NSG_CFG(24,   12,  12,  12,  12,  12,  12)
NSG_CFG(32,   8,   8,   8,   8,   8,   8)
NSG_CFG(48,   12,  12,  12,  12,  12,  12)
NSG_CFG(64,   8,   8,   8,   8,   8,   8)
NSG_CFG(96,   12,  12,  12,  12,  12,  12)
NSG_CFG(128,  8,   8,   8,   8,   8,   8)
NSG_CFG(256,  8,   8,   8,   8,   4,   8)
NSG_CFG(512,  8,   8,   8,   16,  4,   8)
NSG_CFG(1024, 16,  16,  16,  16,  8,   8)
#undef NSG_CFG

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

struct nsg_ctx {
    nsg_params p;
    int N = 0, kmax = 0, NK = 0, NZP = 0, nspect = 0;
    int ncu = 256, zblocks_per_cu = 8;  // persistent z-pass grid = ncu * zblocks_per_cu blocks
    int zocc = 1;                       // >= 3: z kernel compiled for 3 waves/SIMD (fewer VGPRs)
    bool fp32 = false;
    hipStream_t s = nullptr;
    double2* uh = nullptr;   // state [3][NK][NK][NZP]
    double2* aux = nullptr;  // second FP64 state (nsg_rhs input)
    void* K[3] = {nullptr, nullptr, nullptr};
    void* B1 = nullptr;
    void* B2 = nullptr;
    double2 *twN64 = nullptr, *twM64 = nullptr;
    float2 *twN32 = nullptr, *twM32 = nullptr;
    DevScalars* sc = nullptr;
    double* fslots = nullptr;
    double* part = nullptr;
    size_t part_n = 0;
    std::string err;
    // profiling
    bool prof = false;
    hipEvent_t ev[2] = {nullptr, nullptr};
    double prof_ms[5] = {0, 0, 0, 0, 0};
    long prof_cnt[5] = {0, 0, 0, 0, 0};
};

static int fail(nsg_ctx* c, int code, const char* fmt, ...) __attribute__((format(printf, 3, 4)));
static int fail(nsg_ctx* c, int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    return code;
}

#define HIPCHK(ctx, expr)                                                                        \
    do {                                                                                         \
        hipError_t e_ = (expr);                                                                  \
        if (e_ != hipSuccess)                                                                    \
            return fail((ctx), NSG_ERR_DEVICE, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                        __FILE__, __LINE__);                                                     \
    } while (0)

#define LAUNCHCHK(ctx)                                                                            \
    do {                                                                                          \
        hipError_t e_ = hipGetLastError();                                                        \
        if (e_ != hipSuccess)                                                                     \
            return fail((ctx), NSG_ERR_DEVICE, "kernel launch failed: %s (%s:%d)", hipGetErrorString(e_), \
                        __FILE__, __LINE__);                                                      \
    } while (0)

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
// pass launchers
// ------------------------------------------------------------------------------------------
template <typename T> static const void* twN_of(nsg_ctx* c) {
    if constexpr (std::is_same<T, double>::value) return c->twN64; else return c->twN32;
}
template <typename T> static const void* twM_of(nsg_ctx* c) {
    if constexpr (std::is_same<T, double>::value) return c->twM64; else return c->twM32;
}
template <typename T> static constexpr bool is64() { return std::is_same<T, double>::value; }
template <typename T, int N> static constexpr int WXof() { return is64<T>() ? Cfg<N>::WX64 : Cfg<N>::WX32; }
template <typename T, int N> static constexpr int WYof() { return is64<T>() ? Cfg<N>::WY64 : Cfg<N>::WY32; }
template <typename T, int N> static constexpr int EXof() { return is64<T>() ? Cfg<N>::EX64 : Cfg<N>::EX32; }
template <typename T, int N> static constexpr int EYof() { return is64<T>() ? Cfg<N>::EY64 : Cfg<N>::EY32; }
template <typename T, int N> static constexpr int EZof() { return is64<T>() ? Cfg<N>::EZ64 : Cfg<N>::EZ32; }

// grid of the x passes: 2-D (tiles, rows), or 1-D XCD-paired when a tile row is under 128 bytes
template <typename T, int W> static dim3 xgrid(int ntiles, int nrows) {
    if (W * sizeof(cpx_t<T>) >= 128) return dim3(ntiles, nrows);
    const long long pairs = (long long)(ntiles / 2) * nrows;
    return dim3((unsigned)(16 * ((pairs + 7) / 8)));
}

template <typename T, int N>
static int launch_pass1(nsg_ctx* c, const double2* uh, const void* kprev, int stage, int sel, int force) {
    constexpr int EPT = EXof<T, N>(), W = WXof<T, N>(), TPF = N / EPT;
    static_assert(TPF * W <= 1024, "block too large");
    P1Args a;
    a.uh = uh; a.kprev = kprev; a.out = c->B1; a.twN = twN_of<T>(c); a.sc = c->sc; a.fslots = c->fslots;
    a.nu = c->p.nu; a.kf2 = c->p.force_kf * c->p.force_kf;
    a.visc_pow = c->p.visc_pow; a.scheme = c->p.scheme; a.stage = stage; a.force = force; a.sel = sel;
    a.kmax = c->kmax; a.NK = c->NK; a.NZP = c->NZP;
    const dim3 grid = xgrid<T, W>(c->NZP / W, c->NK);
    prof_begin(c);
    hipLaunchKernelGGL((pass1_kernel<T, N, EPT, W>), grid, dim3(TPF * W), 0, c->s, a);
    LAUNCHCHK(c);
    prof_end(c, 0);
    return NSG_OK;
}

template <typename T, int N, int DIR>
static int launch_passy(nsg_ctx* c, const void* in, void* out, int F) {
    constexpr int EPT = EYof<T, N>(), W = WYof<T, N>(), TPF = N / EPT;
    static_assert(TPF * W <= 1024, "block too large");
    PYArgs a;
    a.in = in; a.out = out; a.twN = twN_of<T>(c); a.kmax = c->kmax; a.NK = c->NK; a.NZP = c->NZP;
    prof_begin(c);
    hipLaunchKernelGGL((passy_kernel<T, N, EPT, W, DIR>), dim3(c->NZP / W, N, F), dim3(TPF * W), 0, c->s, a);
    LAUNCHCHK(c);
    prof_end(c, DIR > 0 ? 1 : 3);
    return NSG_OK;
}

// RHS z pass: persistent wavefronts (WPB per block) when a line fits in a wavefront
template <typename T, int N> static constexpr bool zrhs_persistent() {
    constexpr int TPFZ = N / 2 / EZof<T, N>();
    return TPFZ <= 64 && 64 % TPFZ == 0 && ((N / 2) & (N / 2 - 1)) == 0;
}

template <typename T, int N>
static int launch_passz_rhs(nsg_ctx* c, const void* in, void* out, double scale, int track_max) {
    constexpr int EPTZ = EZof<T, N>(), TPFZ = N / 2 / EPTZ, LPW = 64 / TPFZ, WPB = 4;
    static_assert(EPTZ >= 1 && LPW >= 1, "bad z configuration");
    PZArgs a;
    a.in = in; a.out = out; a.twM = twM_of<T>(c); a.twN = twN_of<T>(c); a.sc = c->sc;
    a.scale = scale; a.kmax = c->kmax; a.NZP = c->NZP; a.track_max = track_max;
    const long long groups = ((long long)N * N + LPW - 1) / LPW;
    long long blocks = (groups + WPB - 1) / WPB;
    const long long cap = (long long)c->ncu * c->zblocks_per_cu;
    if (blocks > cap) blocks = cap;
    prof_begin(c);
    if (c->zocc >= 3)
        hipLaunchKernelGGL((passz_rhs_kernel<T, N, EPTZ, WPB, 3>), dim3((unsigned)blocks), dim3(64 * WPB), 0, c->s, a);
    else
        hipLaunchKernelGGL((passz_rhs_kernel<T, N, EPTZ, WPB, 1>), dim3((unsigned)blocks), dim3(64 * WPB), 0, c->s, a);
    LAUNCHCHK(c);
    prof_end(c, 2);
    return NSG_OK;
}

template <typename T, int N, int MODE>
static int launch_passz(nsg_ctx* c, const void* in, void* out, double scale, int track_max) {
    constexpr int EPTZ = EZof<T, N>(), TPFZ = N / 2 / EPTZ;
    constexpr int LPB = (256 / TPFZ) > 0 ? (256 / TPFZ) : 1;
    PZArgs a;
    a.in = in; a.out = out; a.twM = twM_of<T>(c); a.twN = twN_of<T>(c); a.sc = c->sc;
    a.scale = scale; a.kmax = c->kmax; a.NZP = c->NZP; a.track_max = track_max;
    const long long lines = (long long)N * N;
    const unsigned blocks = (unsigned)((lines + LPB - 1) / LPB);
    prof_begin(c);
    hipLaunchKernelGGL((passz_kernel<T, N, EPTZ, LPB, MODE>), dim3(blocks), dim3(TPFZ * LPB), 0, c->s, a);
    LAUNCHCHK(c);
    prof_end(c, 2);
    return NSG_OK;
}

template <typename T, int N, int MODE>
static int launch_pass5(nsg_ctx* c, double2* uh, int stage, void* kout, int force, double norm) {
    constexpr int EPT = EXof<T, N>(), W = WXof<T, N>(), TPF = N / EPT;
    static_assert(TPF * W <= 1024, "block too large");
    P5Args a;
    a.in = c->B2; a.uh = uh; a.k1 = c->K[0]; a.k2 = c->K[1]; a.k3 = c->K[2]; a.kout = kout;
    a.twN = twN_of<T>(c); a.sc = c->sc; a.fslots = c->fslots;
    a.nu = c->p.nu; a.kf2 = c->p.force_kf * c->p.force_kf; a.eps_in = c->p.force_eps; a.norm = norm;
    a.visc_pow = c->p.visc_pow; a.scheme = c->p.scheme; a.stage = stage; a.force = force;
    a.kmax = c->kmax; a.NK = c->NK; a.NZP = c->NZP;
    const dim3 grid = xgrid<T, W>(c->NZP / W, c->NK);
    prof_begin(c);
    hipLaunchKernelGGL((pass5_kernel<T, N, EPT, W, MODE>), grid, dim3(TPF * W), 0, c->s, a);
    LAUNCHCHK(c);
    prof_end(c, 4);
    return NSG_OK;
}

// one RK stage s (1..4) of the nonlinear term on state `uh`, writing K_s (s<4) or the update (s=4)
template <typename T, int N>
static int run_stage(nsg_ctx* c, double2* uh, int stage, bool track_max, bool final_update) {
    const int force = c->p.force_eps > 0.0 ? 1 : 0;
    const void* kprev = stage > 1 ? c->K[stage - 2] : nullptr;
    const double N3 = (double)N * N * N;
    int rc;
    if ((rc = launch_pass1<T, N>(c, uh, kprev, stage, 0, force))) return rc;
    if ((rc = launch_passy<T, N, +1>(c, c->B1, c->B2, 6))) return rc;
    if constexpr (zrhs_persistent<T, N>()) {
        if ((rc = launch_passz_rhs<T, N>(c, c->B2, c->B1, 1.0 / N3, track_max ? 1 : 0))) return rc;
    } else {
        if ((rc = launch_passz<T, N, ZM_RHS>(c, c->B2, c->B1, 1.0 / N3, track_max ? 1 : 0))) return rc;
    }
    if (track_max && c->p.cfl > 0.0) {
        hipLaunchKernelGGL(cfl_kernel, dim3(1), dim3(1), 0, c->s, c->sc);
        LAUNCHCHK(c);
    }
    if ((rc = launch_passy<T, N, -1>(c, c->B1, c->B2, 3))) return rc;
    const int st = final_update ? 4 : (stage < 4 ? stage : 1);
    void* kout = (st < 4) ? c->K[st - 1] : nullptr;
    return launch_pass5<T, N, EP_STAGE>(c, uh, st, kout, force, 1.0 / (N3 * N3));
}

template <typename T, int N> static int run_step(nsg_ctx* c) {
    int rc;
    hipLaunchKernelGGL(step_begin_kernel, dim3(1), dim3(1), 0, c->s, c->sc);
    LAUNCHCHK(c);
    for (int s = 1; s <= 4; ++s)
        if ((rc = run_stage<T, N>(c, c->uh, s, s == 1, s == 4))) return rc;
    hipLaunchKernelGGL(step_end_kernel, dim3(1), dim3(1), 0, c->s, c->sc);
    LAUNCHCHK(c);
    return NSG_OK;
}

template <int N> static int run_steps_N(nsg_ctx* c, long n) {
    for (long i = 0; i < n; ++i) {
        int rc = c->fp32 ? run_step<float, N>(c) : run_step<double, N>(c);
        if (rc) return rc;
    }
    return NSG_OK;
}

// ------------------------------------------------------------------------------------------
// C ABI (C linkage comes from the declarations in nsg.h)
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

static void free_all(nsg_ctx* c) {
    if (!c) return;
    void* bufs[] = {c->uh, c->aux, c->K[0], c->K[1], c->K[2], c->B1, c->B2, c->twN64, c->twM64,
                    c->twN32, c->twM32, c->sc, c->fslots, c->part};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    if (c->ev[0]) (void)hipEventDestroy(c->ev[0]);
    if (c->ev[1]) (void)hipEventDestroy(c->ev[1]);
    if (c->s) (void)hipStreamDestroy(c->s);
}

void nsg_destroy(nsg_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->p.device);
    if (ctx->s) (void)hipStreamSynchronize(ctx->s);
    free_all(ctx);
    delete ctx;
}

static thread_local std::string g_create_err;

int nsg_create(const nsg_params* p, nsg_ctx** out) {
    if (!p || !out) return NSG_ERR_INVALID;
    *out = nullptr;
    nsg_ctx* c = new (std::nothrow) nsg_ctx();
    if (!c) return NSG_ERR_OOM;
    c->p = *p;
    const int N = p->N[0];
    auto bad = [&](const char* msg) {
        g_create_err = msg;
        c->err = msg;
        *out = c;  // hand back the context so the caller can read nsg_last_error, then destroy it
        return NSG_ERR_INVALID;
    };
    if (p->N[1] != N || p->N[2] != N) return bad("only cubic grids are supported (N[0]==N[1]==N[2])");
    if (dispatch_N(N, [](auto) { return 0; }) != 0) return bad("unsupported N (8,12,16,24,32,48,64,96,128,256,512,1024)");
    if (!(p->nu >= 0.0) || !(p->dt > 0.0)) return bad("nu must be >= 0 and dt > 0");
    if (p->visc_pow < 1 || p->visc_pow > 4) return bad("visc_pow must be 1..4");
    if (p->scheme != NSG_SCHEME_RK4CN && p->scheme != NSG_SCHEME_IFRK4) return bad("unknown scheme");
    if (p->precision != NSG_FP64 && p->precision != NSG_FP32BULK) return bad("unknown precision");
    if (p->force_eps > 0.0 && !(p->force_kf > 0.0 && p->force_kf < 7.5)) return bad("force_kf must be in (0, 7.5)");
    c->N = N;
    c->kmax = N / 3;
    c->NK = 2 * c->kmax + 1;
    c->NZP = ((c->kmax + 1 + 15) / 16) * 16;
    c->nspect = (int)std::sqrt(3.0 * (N / 2.0) * (N / 2.0)) + 1;
    c->fp32 = p->precision == NSG_FP32BULK;

    hipError_t e = hipSetDevice(p->device);
    if (e != hipSuccess) { *out = c; return fail(c, NSG_ERR_DEVICE, "hipSetDevice(%d): %s", p->device, hipGetErrorString(e)); }
    *out = c;
    HIPCHK(c, hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking));
    {
        hipDeviceProp_t prop;
        HIPCHK(c, hipGetDeviceProperties(&prop, p->device));
        c->ncu = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
        const char* zb = std::getenv("NSG_ZBLOCKS_PER_CU");
        if (zb && std::atoi(zb) > 0) c->zblocks_per_cu = std::atoi(zb);
        const char* zo = std::getenv("NSG_ZOCC");
        if (zo && std::atoi(zo) > 0) c->zocc = std::atoi(zo);
    }
    HIPCHK(c, hipEventCreate(&c->ev[0]));
    HIPCHK(c, hipEventCreate(&c->ev[1]));

    const size_t NK = c->NK, NZP = c->NZP, Nl = N;
    const size_t nstate = 3 * NK * NK * NZP;
    const size_t ebulk = c->fp32 ? sizeof(float2) : sizeof(double2);
    // B1: A1 [6][N][NK][NZP], A3 [3][N][N][NZP], real AoS [N^3][3]  (FP64 sized: IC/realspace paths run FP64)
    size_t b1 = std::max(6 * Nl * (size_t)xplane(c->NK, c->NZP), 3 * Nl * Nl * NZP) * sizeof(double2);
    b1 = std::max(b1, 3 * Nl * Nl * Nl * sizeof(double));
    // B2: A2 [6][N][N][NZP], A4, full AoS spectral staging [N][N][N/2+1][3], real AoS
    size_t b2 = std::max(6 * Nl * Nl * NZP, 3 * Nl * (size_t)xplane(c->NK, c->NZP)) * sizeof(double2);
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
    for (int i = 0; i < 3; ++i)
        if ((rc = alloc(&c->K[i], nstate * ebulk))) return rc;
    if ((rc = alloc(&c->B1, b1))) return rc;
    if ((rc = alloc(&c->B2, b2))) return rc;
    if ((rc = alloc((void**)&c->sc, sizeof(DevScalars)))) return rc;
    if ((rc = alloc((void**)&c->fslots, sizeof(double) * NK))) return rc;
    c->part_n = std::max<size_t>(7 * (size_t)c->nspect, (nstate / 3 + 255) / 256 + 1);
    if ((rc = alloc((void**)&c->part, sizeof(double) * c->part_n))) return rc;
    if ((rc = upload_table(c, N, &c->twN64, &c->twN32))) return rc;
    if ((rc = upload_table(c, N / 2, &c->twM64, &c->twM32))) return rc;
    HIPCHK(c, hipMemset(c->uh, 0, nstate * sizeof(double2)));
    HIPCHK(c, hipMemset(c->fslots, 0, sizeof(double) * NK));
    for (int i = 0; i < 3; ++i) HIPCHK(c, hipMemset(c->K[i], 0, nstate * ebulk));
    DevScalars hs{};
    hs.dt = p->dt; hs.t = p->t0; hs.maxbits = 0; hs.max_last = 0.0; hs.steps = 0;
    hs.cfl = p->cfl; hs.dx = 2.0 * M_PI / N;
    HIPCHK(c, hipMemcpy(c->sc, &hs, sizeof hs, hipMemcpyHostToDevice));
    HIPCHK(c, hipDeviceSynchronize());  // the null-stream memsets above must land before work on c->s
    return NSG_OK;
}

int nsg_set_state(nsg_ctx* c, const double* u_hat) {
    if (!c || !u_hat) return NSG_ERR_INVALID;
    HIPCHK(c, hipSetDevice(c->p.device));
    const size_t full = 3 * (size_t)c->N * c->N * (c->N / 2 + 1) * sizeof(double2);
    HIPCHK(c, hipMemcpyAsync(c->B2, u_hat, full, hipMemcpyHostToDevice, c->s));
    hipLaunchKernelGGL(aos_to_state_kernel, dim3(2048), dim3(256), 0, c->s, (const double2*)c->B2, c->uh, c->N, c->kmax, c->NK, c->NZP);
    LAUNCHCHK(c);
    HIPCHK(c, hipStreamSynchronize(c->s));
    return NSG_OK;
}

static int download_compact(nsg_ctx* c, const void* src, bool bulk, double* out) {
    const size_t full = 3 * (size_t)c->N * c->N * (c->N / 2 + 1) * sizeof(double2);
    if (bulk && c->fp32)
        hipLaunchKernelGGL(compact_to_aos_kernel<float2>, dim3(2048), dim3(256), 0, c->s, (const float2*)src, (double2*)c->B2, c->N, c->kmax, c->NK, c->NZP);
    else
        hipLaunchKernelGGL(compact_to_aos_kernel<double2>, dim3(2048), dim3(256), 0, c->s, (const double2*)src, (double2*)c->B2, c->N, c->kmax, c->NK, c->NZP);
    LAUNCHCHK(c);
    HIPCHK(c, hipMemcpyAsync(out, c->B2, full, hipMemcpyDeviceToHost, c->s));
    HIPCHK(c, hipStreamSynchronize(c->s));
    return NSG_OK;
}

int nsg_get_state(nsg_ctx* c, double* u_hat) {
    if (!c || !u_hat) return NSG_ERR_INVALID;
    HIPCHK(c, hipSetDevice(c->p.device));
    return download_compact(c, c->uh, false, u_hat);
}

// r2c of a real AoS field already in B2 into the state (FP64 path)
template <int N> static int r2c_into_state(nsg_ctx* c) {
    int rc;
    if ((rc = launch_passz<double, N, ZM_R2C>(c, c->B2, c->B1, 1.0, 0))) return rc;
    if ((rc = launch_passy<double, N, -1>(c, c->B1, c->B2, 3))) return rc;
    return launch_pass5<double, N, EP_STORE>(c, c->uh, 1, nullptr, 0, 1.0);
}

int nsg_set_state_real(nsg_ctx* c, const double* u) {
    if (!c || !u) return NSG_ERR_INVALID;
    HIPCHK(c, hipSetDevice(c->p.device));
    const size_t bytes = 3 * (size_t)c->N * c->N * c->N * sizeof(double);
    HIPCHK(c, hipMemcpyAsync(c->B2, u, bytes, hipMemcpyHostToDevice, c->s));
    int rc = dispatch_N(c->N, [&](auto n) { return r2c_into_state<decltype(n)::value>(c); });
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->s));
    return NSG_OK;
}

int nsg_init_state(nsg_ctx* c, int kind, double k0, double energy, uint64_t seed) {
    if (!c) return NSG_ERR_INVALID;
    HIPCHK(c, hipSetDevice(c->p.device));
    if (kind == NSG_IC_TAYLOR_GREEN || kind == NSG_IC_SHAPIRO) {
        hipLaunchKernelGGL(ic_real_kernel, dim3(4096), dim3(256), 0, c->s, (double*)c->B2, c->N,
                           kind == NSG_IC_TAYLOR_GREEN ? 0 : 1, c->p.nu, 0.0);
        LAUNCHCHK(c);
        int rc = dispatch_N(c->N, [&](auto n) { return r2c_into_state<decltype(n)::value>(c); });
        if (rc) return rc;
    } else if (kind == NSG_IC_K4) {
        if (!(k0 > 0.0) || !(energy >= 0.0)) return fail(c, NSG_ERR_INVALID, "k4 IC needs k0 > 0 and energy >= 0");
        const long long plane = (long long)c->NK * c->NK * c->NZP;
        const unsigned nb = (unsigned)((plane + 255) / 256);
        if ((size_t)nb + 1 > c->part_n) return fail(c, NSG_ERR_INVALID, "internal: partial buffer too small");
        hipLaunchKernelGGL(ic_k4_kernel, dim3(nb), dim3(256), 0, c->s, c->uh, c->part, c->N, c->kmax, c->NK, c->NZP, k0, (unsigned long long)seed);
        LAUNCHCHK(c);
        hipLaunchKernelGGL(sum_parts_kernel, dim3(1), dim3(1), 0, c->s, c->part, (int)nb, c->part + nb);
        LAUNCHCHK(c);
        const double N3 = (double)c->N * c->N * c->N;
        hipLaunchKernelGGL(scale_state_kernel, dim3(2048), dim3(256), 0, c->s, c->uh, 3 * plane, (const double*)(c->part + nb), energy, N3);
        LAUNCHCHK(c);
    } else {
        return fail(c, NSG_ERR_INVALID, "unknown initial condition %d", kind);
    }
    HIPCHK(c, hipStreamSynchronize(c->s));
    return NSG_OK;
}

int nsg_rhs(nsg_ctx* c, const double* u_hat, double* out) {
    if (!c || !u_hat || !out) return NSG_ERR_INVALID;
    HIPCHK(c, hipSetDevice(c->p.device));
    const size_t full = 3 * (size_t)c->N * c->N * (c->N / 2 + 1) * sizeof(double2);
    HIPCHK(c, hipMemcpyAsync(c->B2, u_hat, full, hipMemcpyHostToDevice, c->s));
    hipLaunchKernelGGL(aos_to_state_kernel, dim3(2048), dim3(256), 0, c->s, (const double2*)c->B2, c->aux, c->N, c->kmax, c->NK, c->NZP);
    LAUNCHCHK(c);
    int rc = dispatch_N(c->N, [&](auto n) {
        constexpr int NN = decltype(n)::value;
        return c->fp32 ? run_stage<float, NN>(c, c->aux, 1, false, false) : run_stage<double, NN>(c, c->aux, 1, false, false);
    });
    if (rc) return rc;
    return download_compact(c, c->K[0], true, out);
}

int nsg_step(nsg_ctx* c, long nsteps) {
    if (!c || nsteps < 0) return NSG_ERR_INVALID;
    HIPCHK(c, hipSetDevice(c->p.device));
    int rc = dispatch_N(c->N, [&](auto n) { return run_steps_N<decltype(n)::value>(c, nsteps); });
    if (rc) return rc;
    DevScalars hs;
    HIPCHK(c, hipMemcpyAsync(&hs, c->sc, sizeof hs, hipMemcpyDeviceToHost, c->s));
    HIPCHK(c, hipStreamSynchronize(c->s));
    if (nsteps > 0) {
        if (!std::isfinite(hs.max_last) || !std::isfinite(hs.dt))
            return fail(c, NSG_ERR_NONFINITE, "non-finite velocity after step %lld", hs.steps);
        if (c->p.max_speed > 0.0 && hs.max_last > c->p.max_speed)
            return fail(c, NSG_ERR_BLOWUP, "max|u| = %g exceeds the limit %g at step %lld", hs.max_last, c->p.max_speed, hs.steps);
    }
    return NSG_OK;
}

int nsg_measures(nsg_ctx* c, nsg_totals* out, double* E_k, double* Ens_k, int n_spect, int quirk) {
    if (!c || !out) return NSG_ERR_INVALID;
    if ((E_k || Ens_k) && n_spect < c->nspect) return fail(c, NSG_ERR_INVALID, "n_spect %d < %d", n_spect, c->nspect);
    HIPCHK(c, hipSetDevice(c->p.device));
    const int ns = c->nspect;
    hipLaunchKernelGGL(measures_kernel, dim3(ns), dim3(256), 0, c->s, c->uh, c->part, c->N, c->kmax, c->NK, c->NZP,
                       c->p.nu, c->p.visc_pow, quirk ? 1 : 0);
    LAUNCHCHK(c);
    std::vector<double> h((size_t)ns * 7);
    HIPCHK(c, hipMemcpyAsync(h.data(), c->part, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->s));
    HIPCHK(c, hipStreamSynchronize(c->s));
    const double N3 = (double)c->N * c->N * c->N;
    const double nf = 0.5 / (N3 * N3), cf = 8.0 * M_PI * M_PI * M_PI;
    double t[5] = {0, 0, 0, 0, 0};
    for (int b = 0; b < ns; ++b)
        for (int j = 0; j < 5; ++j) t[j] += h[(size_t)b * 7 + j];
    out->energy = t[0] * cf * nf;
    out->enstrophy = t[1] * cf * nf;
    out->palinstrophy = t[2] * cf * nf;
    out->helicity = t[3] * cf * nf;
    out->dissipation = t[4] * 2.0 * cf * nf;
    for (int b = 0; b < n_spect && (E_k || Ens_k); ++b) {
        if (E_k) E_k[b] = b < ns ? cf * nf * h[(size_t)b * 7 + 5] : 0.0;
        if (Ens_k) Ens_k[b] = b < ns ? cf * nf * h[(size_t)b * 7 + 6] : 0.0;
    }
    const double* vals[] = {&out->energy, &out->enstrophy, &out->palinstrophy, &out->helicity, &out->dissipation};
    for (const double* v : vals)
        if (!std::isfinite(*v)) return fail(c, NSG_ERR_NONFINITE, "non-finite diagnostics");
    return NSG_OK;
}

template <int N> static int realspace_N(nsg_ctx* c, int which) {
    int rc;
    if ((rc = launch_pass1<double, N>(c, c->uh, nullptr, 1, which == 0 ? 1 : 2, 0))) return rc;
    // pass 2 on three fields: A1 [3][N][NK][NZP] -> A2 [3][N][N][NZP]
    if ((rc = launch_passy<double, N, +1>(c, c->B1, c->B2, 3))) return rc;
    const double N3 = (double)N * N * N;
    return launch_passz<double, N, ZM_REAL>(c, c->B2, c->B1, 1.0 / N3, 0);
}

int nsg_realspace(nsg_ctx* c, int which, double* out) {
    if (!c || !out || (which != 0 && which != 1)) return NSG_ERR_INVALID;
    HIPCHK(c, hipSetDevice(c->p.device));
    int rc = dispatch_N(c->N, [&](auto n) { return realspace_N<decltype(n)::value>(c, which); });
    if (rc) return rc;
    const size_t bytes = 3 * (size_t)c->N * c->N * c->N * sizeof(double);
    HIPCHK(c, hipMemcpyAsync(out, c->B1, bytes, hipMemcpyDeviceToHost, c->s));
    HIPCHK(c, hipStreamSynchronize(c->s));
    return NSG_OK;
}

int nsg_get_time(nsg_ctx* c, double* t, double* dt, double* max_speed) {
    if (!c) return NSG_ERR_INVALID;
    HIPCHK(c, hipSetDevice(c->p.device));
    DevScalars hs;
    HIPCHK(c, hipMemcpyAsync(&hs, c->sc, sizeof hs, hipMemcpyDeviceToHost, c->s));
    HIPCHK(c, hipStreamSynchronize(c->s));
    if (t) *t = hs.t;
    if (dt) *dt = hs.dt;
    if (max_speed) *max_speed = hs.max_last;
    return NSG_OK;
}

int nsg_set_time(nsg_ctx* c, double t, double dt) {
    if (!c || !(dt > 0.0)) return NSG_ERR_INVALID;
    HIPCHK(c, hipSetDevice(c->p.device));
    DevScalars hs;
    HIPCHK(c, hipMemcpyAsync(&hs, c->sc, sizeof hs, hipMemcpyDeviceToHost, c->s));
    HIPCHK(c, hipStreamSynchronize(c->s));
    hs.t = t;
    hs.dt = dt;
    HIPCHK(c, hipMemcpyAsync(c->sc, &hs, sizeof hs, hipMemcpyHostToDevice, c->s));
    HIPCHK(c, hipStreamSynchronize(c->s));
    return NSG_OK;
}

static void pass_bytes(const nsg_ctx* c, double* b /*[5]: stage-averaged*/) {
    const double N = c->N, NK = c->NK, NZL = c->kmax + 1;
    const double E = c->fp32 ? 8.0 : 16.0, S = 16.0;
    const double st = 3 * NK * NK * NZL;  // complex elements of a 3-vector state
    // pass 1: read state (+ K_{s-1} for s > 1), write 6 fields
    b[0] = st * S + 0.75 * st * E + 6 * N * NK * NZL * E;
    b[1] = (6 * N * NK * NZL + 6 * N * N * NZL) * E;
    b[2] = (6 * N * N * NZL + 3 * N * N * NZL) * E;
    b[3] = (3 * N * N * NZL + 3 * N * NK * NZL) * E;
    // pass 5: read A4, write K_s (s<4) | read state + K1..K3, write state (s=4)
    b[4] = 3 * N * NK * NZL * E + 0.75 * st * E + 0.25 * (st * S + 3 * st * E + st * S);
}

double nsg_step_bytes(const nsg_ctx* c) {
    if (!c) return 0.0;
    double b[5];
    pass_bytes(c, b);
    return 4.0 * (b[0] + b[1] + b[2] + b[3] + b[4]);
}

int nsg_profile_passes(nsg_ctx* c, int iters, double* ms, double* bytes) {
    if (!c || iters <= 0 || !ms) return NSG_ERR_INVALID;
    HIPCHK(c, hipSetDevice(c->p.device));
    for (int k = 0; k < 5; ++k) { c->prof_ms[k] = 0.0; c->prof_cnt[k] = 0; }
    HIPCHK(c, hipStreamSynchronize(c->s));
    c->prof = true;
    int rc = dispatch_N(c->N, [&](auto n) { return run_steps_N<decltype(n)::value>(c, iters); });
    c->prof = false;
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->s));
    double tot = 0.0;
    for (int k = 0; k < 5; ++k) {
        ms[k] = c->prof_cnt[k] ? c->prof_ms[k] / c->prof_cnt[k] : 0.0;
        tot += c->prof_ms[k];
    }
    ms[5] = tot / iters;
    if (bytes) pass_bytes(c, bytes);
    return NSG_OK;
}
