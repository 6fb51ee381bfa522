// nsg_internal.hpp -- the context of a batch of boxes and helpers shared by the library's
// translation units (nsg_api.hip, nsg_closure.hip); not part of the C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/nsg.h"

namespace nsg {
struct DevScalars;
}

struct nsg_ctx {
    nsg_params p;  // parameters of box 0; the ones a batch shares apply to every box
    int N = 0, kmax = 0, NK = 0, NZP = 0, nspect = 0, nbox = 1;
    int sched = 0;
    bool fp32 = false, any_cfl = false;
    int ncu = 256;
    int zblocks_env = 0;  // NSG_ZBLOCKS_PER_CU: overrides the schedule's persistent z-pass grid
    hipStream_t s = nullptr;
    // per-box element strides of the state/K arrays and of A1..A4
    long long sbox = 0, a1box = 0, a2box = 0, a3box = 0, a4box = 0;
    double2* uh = nullptr;   // states [nbox][3][NK][NK][NZP]
    double2* aux = nullptr;  // a second FP64 state per box (nsg_rhs input, autotuner backup)
    void* K[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};  // stored stage slopes
    int nslopes = 3;                  // K buffers in use: 3 (RK4 schemes), 5 (RK5-CN), 6 (Dormand-Prince)
    int nstage = 4;                   // RHS evaluations per step (try)
    double2* ucand = nullptr;         // Dormand-Prince candidate state
    double* dperr = nullptr;          // Dormand-Prince per-block error sums
    int dp_nslots = 0;                // blocks of the final pass 5 (one error slot each)
    long long dp_rejected = 0;        // Dormand-Prince tries rejected so far
    void* B1 = nullptr;  // A1 / A3 of every box, or one box's real AoS field
    void* B2 = nullptr;  // A2 / A4 of every box, or one box's AoS staging
    double2 *twN64 = nullptr, *twM64 = nullptr;
    float2 *twN32 = nullptr, *twM32 = nullptr;
    nsg::DevScalars* sc = nullptr;  // [nbox]
    double* fslots = nullptr;  // [nbox][NK]
    double* part = nullptr;
    size_t part_n = 0;
    std::vector<double> max_speed;  // per-box blow-up limit (0 = none)
    std::string err;
    // profiling
    bool prof = false;
    hipEvent_t ev[2] = {nullptr, nullptr};
    double prof_ms[5] = {0, 0, 0, 0, 0};
    long prof_cnt[5] = {0, 0, 0, 0, 0};
    // learned closure (single-box contexts): network, its spectral term -P div(tau) added to
    // every stage's nonlinear term (compact FP64 [3][NK][NK][NZP]), scratch of its evaluation
    nsg_cnn* cnn = nullptr;
    double2* closure = nullptr;
    double* cl_real = nullptr;    // 5 real AoS fields [N^3][3]: 3 gradient groups, tau diagonal, tau off-diagonal
    double2* cl_spec = nullptr;   // 2 compact FP64 3-vectors
    double* cl_gain = nullptr;    // per-shell gains of the closure term (kmax + 1 shells), or null
};

static int fail(nsg_ctx* c, int code, const char* fmt, ...) __attribute__((format(printf, 3, 4), unused));
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

// internal helpers implemented in nsg_api.hip
// c2r of one box's compact FP64 3-vector src to a physical real AoS field [N^3][3] at dst
// (B1/B2 are scratch; dst must not lie in B2)
int nsgi_c2r(nsg_ctx* c, const double2* src, double* dst);
// r2c of a physical real AoS field [N^3][3] at src (outside B1) into compact FP64 dst, 2/3 mask applied
int nsgi_r2c(nsg_ctx* c, const double* src, double2* dst);
// real-space velocity of box 0 (1/N^3-normalised c2r; AoS [N^3][3], or three planes [3][N^3]
// when soa) left in the context's scratch B1; *field points at it until the context's next
// operation
int nsgi_realspace(nsg_ctx* c, const double** field, int soa);
// implemented in nsg_noalloc.hip: per-device set-up of nsg_measures_noalloc's static scratch (the
// device address of that array), once per device, with `device` current
void nsgi_noalloc_setup(int device);
// implemented in nsg_closure.hip: evaluate the closure term of box 0's state into c->closure
int nsgi_closure_eval(nsg_ctx* c);
void nsgi_closure_free(nsg_ctx* c);
