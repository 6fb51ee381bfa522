// nsg_closure.hip -- LES filter, subgrid-stress training samples and the learned closure term.
//
// Filter: Gaussian G(k) = exp(-k^2 Delta^2 / 24), Delta = pi / k_c, times the sharp cut of the LES
// grid (|k| <= N_les / 3, the LES grid's own 2/3 rule).  In rotational form the filtered
// equations differ from the LES equations by -P div(tau), tau_ij = F(u_i u_j) - T(ub_i ub_j)
// (F: filter, T: sharp cut alone), because u x w = grad(|u|^2 / 2) - div(u u) and P removes
// gradients.  The network maps grad(ub) to the deviatoric part of tau; the LES adds
// M = -P div(tau_model) to the nonlinear term of every RK stage (evaluated once per step).
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "nsg_cnn.hpp"
#include "nsg_internal.hpp"

namespace {

__device__ __forceinline__ int k_of_cidx(int c, int kmax, int NK) { return c <= kmax ? c : c - NK; }
__device__ __forceinline__ int cidx_of_k(int k, int NK) { return k >= 0 ? k : k + NK; }
__device__ __forceinline__ long long sidx(int ix, int iy, int kz, int NK, int NZP) { return ((long long)iy * NK + ix) * NZP + kz; }

__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// dst = i k_j src (3 components), compact layout
__global__ void grad_spec_kernel(const double2* __restrict__ src, double2* __restrict__ dst, int j, int kmax, int NK, int NZP) {
    const long long plane = (long long)NK * NK * NZP;
    for (long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x; n < plane; n += (long long)gridDim.x * blockDim.x) {
        const int kz = (int)(n % NZP), ix = (int)((n / NZP) % NK), iy = (int)(n / ((long long)NZP * NK));
        const double k = j == 0 ? k_of_cidx(ix, kmax, NK) : (j == 1 ? k_of_cidx(iy, kmax, NK) : (double)kz);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double2 v = src[c * plane + n];
            dst[c * plane + n] = make_double2(-k * v.y, k * v.x);
        }
    }
}

// gradient groups R_j [N^3][3] (R_j[p][i] = d_j u_i) -> network input, channel 3 i + j
__global__ void pack_grad_bf16_kernel(const double* __restrict__ R, long long n, uint16_t* __restrict__ out, float inv) {
    for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < n; p += (long long)gridDim.x * blockDim.x) {
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) out[p * 16 + 3 * i + j] = f2bf((float)(R[(j * n + p) * 3 + i] * inv));
#pragma unroll
        for (int c = 9; c < 16; ++c) out[p * 16 + c] = 0;
    }
}
__global__ void pack_grad_f32_kernel(const double* __restrict__ R, long long n, float* __restrict__ out) {
    for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < n; p += (long long)gridDim.x * blockDim.x)
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) out[p * 9 + 3 * i + j] = (float)R[(j * n + p) * 3 + i];
}

// network output (normalised, 16 ch) -> physical tau: diagonal and off-diagonal AoS [N^3][3]
__global__ void unpack_tau_kernel(const float* __restrict__ o, long long n, double* __restrict__ Td, double* __restrict__ To,
                                  double s0, double s1, double s2, double s3, double s4, double s5) {
    for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < n; p += (long long)gridDim.x * blockDim.x) {
        Td[p * 3 + 0] = o[p * 16 + 0] * s0;
        Td[p * 3 + 1] = o[p * 16 + 1] * s1;
        Td[p * 3 + 2] = o[p * 16 + 2] * s2;
        To[p * 3 + 0] = o[p * 16 + 3] * s3;
        To[p * 3 + 1] = o[p * 16 + 4] * s4;
        To[p * 3 + 2] = o[p * 16 + 5] * s5;
    }
}

// M_i = -i sum_j k_j tau_ij, projected (k = 0 -> 0), 2/3 mask.  Sd = (xx, yy, zz), So = (xy, xz, yz).
// gain (or null): M times gain[b] on the shell b = round(|k|) (k^2 in [b^2 - b + 1, b^2 + b], the
// shells of the energy spectrum)
__global__ void closure_combine_kernel(const double2* __restrict__ Sd, const double2* __restrict__ So, double2* __restrict__ M,
                                       int kmax, int NK, int NZP, const double* __restrict__ gain) {
    const long long plane = (long long)NK * NK * NZP;
    const long long kcut2 = (long long)kmax * kmax;
    for (long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x; n < plane; n += (long long)gridDim.x * blockDim.x) {
        const int kz = (int)(n % NZP), ix = (int)((n / NZP) % NK), iy = (int)(n / ((long long)NZP * NK));
        const long long kx = k_of_cidx(ix, kmax, NK), ky = k_of_cidx(iy, kmax, NK);
        const long long k2 = kx * kx + ky * ky + (long long)kz * kz;
        double2 m[3] = {make_double2(0.0, 0.0), make_double2(0.0, 0.0), make_double2(0.0, 0.0)};
        if (k2 > 0 && k2 <= kcut2 && kz <= kmax) {
            const double2 xx = Sd[n], yy = Sd[plane + n], zz = Sd[2 * plane + n];
            const double2 xy = So[n], xz = So[plane + n], yz = So[2 * plane + n];
            const double a = (double)kx, b = (double)ky, c = (double)kz;
            const double2 d[3] = {make_double2(a * xx.x + b * xy.x + c * xz.x, a * xx.y + b * xy.y + c * xz.y),
                                  make_double2(a * xy.x + b * yy.x + c * yz.x, a * xy.y + b * yy.y + c * yz.y),
                                  make_double2(a * xz.x + b * yz.x + c * zz.x, a * xz.y + b * yz.y + c * zz.y)};
#pragma unroll
            for (int i = 0; i < 3; ++i) m[i] = make_double2(d[i].y, -d[i].x);  // -i d
            const double kinv = 1.0 / (double)k2;
            const double pr = (a * m[0].x + b * m[1].x + c * m[2].x) * kinv, pi = (a * m[0].y + b * m[1].y + c * m[2].y) * kinv;
            m[0].x -= a * pr; m[0].y -= a * pi;
            m[1].x -= b * pr; m[1].y -= b * pi;
            m[2].x -= c * pr; m[2].y -= c * pi;
            if (gain) {
                long long b = (long long)sqrt((double)k2);
                while (b * b > k2) --b;
                while ((b + 1) * (b + 1) <= k2) ++b;
                if (k2 > b * b + b) ++b;  // round(|k|)
                const double g = gain[b];
#pragma unroll
                for (int i = 0; i < 3; ++i) m[i] = make_double2(g * m[i].x, g * m[i].y);
            }
        }
#pragma unroll
        for (int i = 0; i < 3; ++i) M[i * plane + n] = m[i];
    }
}

// LES state from a DNS state: dst(k) = G(k) src(k) (Nl/N)^3 on |k| <= kmax_l (both compact)
__global__ void downsample_kernel(const double2* __restrict__ src, int kmax_s, int NK_s, int NZP_s, double2* __restrict__ dst,
                                  int kmax_d, int NK_d, int NZP_d, double delta2_24, double scale) {
    const long long pd = (long long)NK_d * NK_d * NZP_d, ps = (long long)NK_s * NK_s * NZP_s;
    const long long kcut2 = (long long)kmax_d * kmax_d;
    for (long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x; n < pd; n += (long long)gridDim.x * blockDim.x) {
        const int kz = (int)(n % NZP_d), ix = (int)((n / NZP_d) % NK_d), iy = (int)(n / ((long long)NZP_d * NK_d));
        const int kx = k_of_cidx(ix, kmax_d, NK_d), ky = k_of_cidx(iy, kmax_d, NK_d);
        const long long k2 = (long long)kx * kx + (long long)ky * ky + (long long)kz * kz;
        const bool keep = k2 <= kcut2 && kz <= kmax_d && kz <= kmax_s && kx >= -kmax_s && kx <= kmax_s && ky >= -kmax_s && ky <= kmax_s;
        const double g = keep ? exp(-(double)k2 * delta2_24) * scale : 0.0;
        const long long s = keep ? sidx(cidx_of_k(kx, NK_s), cidx_of_k(ky, NK_s), kz, NK_s, NZP_s) : 0;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double2 v = keep ? src[c * ps + s] : make_double2(0.0, 0.0);
            dst[c * pd + n] = make_double2(g * v.x, g * v.y);
        }
    }
}

// products of a real AoS field u [N^3][3]: P [N^3][3] = (uxux, uyuy, uzuz) (off = 0) or
// (uxuy, uxuz, uyuz) (off = 1)
__global__ void products_kernel(const double* __restrict__ u, double* __restrict__ P, long long n, int off) {
    for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < n; p += (long long)gridDim.x * blockDim.x) {
        const double x = u[p * 3], y = u[p * 3 + 1], z = u[p * 3 + 2];
        if (off) {
            P[p * 3] = x * y; P[p * 3 + 1] = x * z; P[p * 3 + 2] = y * z;
        } else {
            P[p * 3] = x * x; P[p * 3 + 1] = y * y; P[p * 3 + 2] = z * z;
        }
    }
}

__global__ void sub_kernel(const double2* __restrict__ a, const double2* __restrict__ b, double2* __restrict__ o, long long n) {
    for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < n; p += (long long)gridDim.x * blockDim.x)
        o[p] = make_double2(a[p].x - b[p].x, a[p].y - b[p].y);
}

// deviatoric tau as float [N^3][6] from diagonal / off-diagonal AoS fields
__global__ void tau_out_kernel(const double* __restrict__ Td, const double* __restrict__ To, float* __restrict__ out, long long n) {
    for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < n; p += (long long)gridDim.x * blockDim.x) {
        const double tr = (Td[p * 3] + Td[p * 3 + 1] + Td[p * 3 + 2]) / 3.0;
        out[p * 6 + 0] = (float)(Td[p * 3] - tr);
        out[p * 6 + 1] = (float)(Td[p * 3 + 1] - tr);
        out[p * 6 + 2] = (float)(Td[p * 3 + 2] - tr);
        out[p * 6 + 3] = (float)To[p * 3];
        out[p * 6 + 4] = (float)To[p * 3 + 1];
        out[p * 6 + 5] = (float)To[p * 3 + 2];
    }
}

constexpr int GB = 1024, TB = 256;

long long plane_of(const nsg_ctx* c) { return (long long)c->NK * c->NK * c->NZP; }
long long npts(const nsg_ctx* c) { return (long long)c->N * c->N * c->N; }

// R_j (j = 0..2) = d_j u_i of the compact FP64 state `uh` as real AoS fields in R [3][N^3][3]
int gradients(nsg_ctx* c, const double2* uh, double2* tmp, double* R) {
    for (int j = 0; j < 3; ++j) {
        hipLaunchKernelGGL(grad_spec_kernel, dim3(GB), dim3(TB), 0, c->s, uh, tmp, j, c->kmax, c->NK, c->NZP);
        LAUNCHCHK(c);
        int rc = nsgi_c2r(c, tmp, R + (long long)j * npts(c) * 3);
        if (rc) return rc;
    }
    return NSG_OK;
}

int alloc_closure(nsg_ctx* c) {
    if (c->closure) return NSG_OK;
    const long long n = npts(c), pl = plane_of(c);
    hipError_t e = hipMalloc(&c->closure, sizeof(double2) * 3 * pl);
    if (e == hipSuccess) e = hipMalloc(&c->cl_real, sizeof(double) * 5 * 3 * n);
    if (e == hipSuccess) e = hipMalloc(&c->cl_spec, sizeof(double2) * 2 * 3 * pl);
    if (e == hipSuccess) e = hipMemset(c->closure, 0, sizeof(double2) * 3 * pl);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        nsgi_closure_free(c);
        return fail(c, NSG_ERR_OOM, "closure buffers: %s", hipGetErrorString(e));
    }
    return NSG_OK;
}

}  // namespace

void nsgi_closure_free(nsg_ctx* c) {
    if (c->closure) (void)hipFree(c->closure);
    if (c->cl_real) (void)hipFree(c->cl_real);
    if (c->cl_spec) (void)hipFree(c->cl_spec);
    if (c->cl_gain) (void)hipFree(c->cl_gain);
    c->closure = nullptr;
    c->cl_real = nullptr;
    c->cl_spec = nullptr;
    c->cl_gain = nullptr;
    c->cnn = nullptr;
}

int nsgi_closure_eval(nsg_ctx* c) {
    nsg_cnn* net = c->cnn;
    const long long n = npts(c), pl = plane_of(c);
    double* R = c->cl_real;
    double* Td = R + 9 * n;
    double* To = R + 12 * n;
    double2* S0 = c->cl_spec;
    double2* S1 = c->cl_spec + 3 * pl;
    int rc;
    if ((rc = gradients(c, c->uh, S0, R))) return rc;
    hipLaunchKernelGGL(pack_grad_bf16_kernel, dim3(GB), dim3(TB), 0, c->s, (const double*)R, n, net->act[0], (float)(1.0 / net->sig_in));
    LAUNCHCHK(c);
    if ((rc = nsgcnn::forward_dev(net, c->s, 1, false))) return fail(c, rc, "network: %s", net->err.c_str());
    const double* so = net->sig_out;
    hipLaunchKernelGGL(unpack_tau_kernel, dim3(GB), dim3(TB), 0, c->s, (const float*)net->out, n, Td, To, so[0], so[1], so[2], so[3],
                       so[4], so[5]);
    LAUNCHCHK(c);
    if ((rc = nsgi_r2c(c, Td, S0))) return rc;
    if ((rc = nsgi_r2c(c, To, S1))) return rc;
    hipLaunchKernelGGL(closure_combine_kernel, dim3(GB), dim3(TB), 0, c->s, (const double2*)S0, (const double2*)S1, c->closure,
                       c->kmax, c->NK, c->NZP, (const double*)c->cl_gain);
    LAUNCHCHK(c);
    return NSG_OK;
}

// ------------------------------------------------------------------------------------------
// C ABI
// ------------------------------------------------------------------------------------------
int nsg_attach_closure(nsg_ctx* c, nsg_cnn* net) {
    if (!c) return NSG_ERR_INVALID;
    HIPCHK(c, hipSetDevice(c->p.device));
    HIPCHK(c, hipStreamSynchronize(c->s));
    if (!net) {
        nsgi_closure_free(c);
        return NSG_OK;
    }
    if (c->nbox != 1) return fail(c, NSG_ERR_INVALID, "a closure needs a single-box context");
    if (net->N != c->N) return fail(c, NSG_ERR_INVALID, "network grid %d != box grid %d", net->N, c->N);
    if (net->device != c->p.device) return fail(c, NSG_ERR_INVALID, "network and box are on different devices");
    int rc = alloc_closure(c);
    if (rc) return rc;
    c->cnn = net;
    return NSG_OK;
}

int nsg_closure_set_gain(nsg_ctx* c, const double* gain, int n) {
    if (!c) return NSG_ERR_INVALID;
    if (!c->cnn) return fail(c, NSG_ERR_INVALID, "no closure attached");
    HIPCHK(c, hipSetDevice(c->p.device));
    HIPCHK(c, hipStreamSynchronize(c->s));
    if (!gain) {
        if (c->cl_gain) (void)hipFree(c->cl_gain);
        c->cl_gain = nullptr;
        return NSG_OK;
    }
    if (n != c->kmax + 1) return fail(c, NSG_ERR_INVALID, "closure gains: %d shells given, %d expected (0..kmax)", n, c->kmax + 1);
    for (int b = 0; b < n; ++b)
        if (!std::isfinite(gain[b])) return fail(c, NSG_ERR_INVALID, "closure gain of shell %d is not finite", b);
    if (!c->cl_gain) HIPCHK(c, hipMalloc(&c->cl_gain, sizeof(double) * n));
    HIPCHK(c, hipMemcpy(c->cl_gain, gain, sizeof(double) * n, hipMemcpyHostToDevice));
    return NSG_OK;
}

int nsg_closure_term(nsg_ctx* c, double* out) {
    if (!c || !out) return NSG_ERR_INVALID;
    if (!c->cnn) return fail(c, NSG_ERR_INVALID, "no closure attached");
    HIPCHK(c, hipSetDevice(c->p.device));
    int rc = nsgi_closure_eval(c);
    if (rc) return rc;
    // compact -> host AoS through the state download path
    const double2* save = c->uh;
    c->uh = c->closure;
    rc = nsg_get_state(c, out);
    c->uh = const_cast<double2*>(save);
    return rc;
}

int nsg_resample(nsg_ctx* src, nsg_ctx* dst) {
    if (!src || !dst) return NSG_ERR_INVALID;
    if (src->nbox != 1 || dst->nbox != 1) return fail(dst, NSG_ERR_INVALID, "single-box contexts only");
    if (src->p.device != dst->p.device) return fail(dst, NSG_ERR_INVALID, "contexts on different devices");
    HIPCHK(dst, hipSetDevice(dst->p.device));
    HIPCHK(src, hipStreamSynchronize(src->s));
    const double scale = std::pow((double)dst->N / src->N, 3);
    hipLaunchKernelGGL(downsample_kernel, dim3(GB), dim3(TB), 0, dst->s, (const double2*)src->uh, src->kmax, src->NK, src->NZP,
                       dst->uh, dst->kmax, dst->NK, dst->NZP, 0.0, scale);
    LAUNCHCHK(dst);
    HIPCHK(dst, hipStreamSynchronize(dst->s));
    double t = 0.0, dt = 0.0;
    int rc = nsg_get_time(src, &t, &dt, nullptr);
    if (rc) return fail(dst, rc, "%s", src->err.c_str());
    return nsg_set_time(dst, t, dt * src->N / dst->N);
}

int nsg_sgs_sample(nsg_ctx* dns, nsg_ctx* les, double kc, float* x_out, float* tau_out) {
    if (!dns || !les || !(kc > 0.0)) return NSG_ERR_INVALID;
    if (dns->nbox != 1 || les->nbox != 1) return fail(les, NSG_ERR_INVALID, "single-box contexts only");
    if (les->N >= dns->N) return fail(les, NSG_ERR_INVALID, "the LES grid must be coarser than the DNS grid");
    if (dns->p.device != les->p.device) return fail(les, NSG_ERR_INVALID, "contexts on different devices");
    HIPCHK(les, hipSetDevice(les->p.device));
    HIPCHK(dns, hipStreamSynchronize(dns->s));
    const double delta = M_PI / kc;
    const double scale = std::pow((double)les->N / dns->N, 3);
    // filtered, downsampled state -> les->uh
    hipLaunchKernelGGL(downsample_kernel, dim3(GB), dim3(TB), 0, les->s, (const double2*)dns->uh, dns->kmax, dns->NK, dns->NZP,
                       les->uh, les->kmax, les->NK, les->NZP, delta * delta / 24.0, scale);
    LAUNCHCHK(les);
    const long long nl = npts(les), pll = plane_of(les);
    double* Rl = nullptr;       // 5 LES real fields [N^3][3]
    double2* Sl = nullptr;      // 4 LES compact 3-vectors
    double* Ud = nullptr;       // DNS real u and product fields
    double2* Sd = nullptr;      // DNS compact 3-vector
    auto cleanup = [&]() {
        for (void* p : {(void*)Rl, (void*)Sl, (void*)Ud, (void*)Sd})
            if (p) (void)hipFree(p);
    };
    int rc = NSG_OK;
    auto hip = [&](hipError_t e, const char* what) {
        if (e != hipSuccess && rc == NSG_OK) rc = fail(les, NSG_ERR_DEVICE, "%s: %s", what, hipGetErrorString(e));
        return e == hipSuccess;
    };
    if (!hip(hipMalloc(&Rl, sizeof(double) * 5 * 3 * nl), "hipMalloc") ||
        !hip(hipMalloc(&Sl, sizeof(double2) * 4 * 3 * pll), "hipMalloc")) {
        cleanup();
        return rc;
    }
    double2* Fd = Sl;
    double2* Fo = Sl + 3 * pll;
    double2* Qd = Sl + 6 * pll;
    double2* Qo = Sl + 9 * pll;
    if (x_out) {
        if ((rc = gradients(les, les->uh, Fd, Rl))) { cleanup(); return rc; }
        float* xd = (float*)(Rl + 9 * nl);  // tau scratch reused for the float copy
        hipLaunchKernelGGL(pack_grad_f32_kernel, dim3(GB), dim3(TB), 0, les->s, (const double*)Rl, nl, xd);
        if (!hip(hipGetLastError(), "pack") || !hip(hipMemcpyAsync(x_out, xd, sizeof(float) * 9 * nl, hipMemcpyDeviceToHost, les->s), "copy") ||
            !hip(hipStreamSynchronize(les->s), "sync")) {
            cleanup();
            return rc;
        }
    }
    if (tau_out) {
        const long long nd = npts(dns), pld = plane_of(dns);
        if (!hip(hipMalloc(&Ud, sizeof(double) * 2 * 3 * nd), "hipMalloc (DNS real fields)") ||
            !hip(hipMalloc(&Sd, sizeof(double2) * 3 * pld), "hipMalloc (DNS spectral field)")) {
            cleanup();
            return rc;
        }
        double* Pd = Ud + 3 * nd;
        // F(u_i u_j) from the DNS grid
        if ((rc = nsgi_c2r(dns, dns->uh, Ud))) { cleanup(); return fail(les, rc, "DNS c2r: %s", dns->err.c_str()); }
        for (int off = 0; off < 2 && rc == NSG_OK; ++off) {
            hipLaunchKernelGGL(products_kernel, dim3(GB), dim3(TB), 0, dns->s, (const double*)Ud, Pd, nd, off);
            if (!hip(hipGetLastError(), "products")) break;
            if ((rc = nsgi_r2c(dns, Pd, Sd))) { rc = fail(les, rc, "DNS r2c: %s", dns->err.c_str()); break; }
            hipLaunchKernelGGL(downsample_kernel, dim3(GB), dim3(TB), 0, dns->s, (const double2*)Sd, dns->kmax, dns->NK, dns->NZP,
                               off ? Fo : Fd, les->kmax, les->NK, les->NZP, delta * delta / 24.0, scale);
            hip(hipGetLastError(), "downsample");
        }
        if (rc == NSG_OK) hip(hipStreamSynchronize(dns->s), "sync");
        // T(ub_i ub_j) on the LES grid (alias-free for |k| <= N_les / 3)
        double* Ub = Rl;
        double* Pb = Rl + 3 * nl;
        if (rc == NSG_OK && (rc = nsgi_c2r(les, les->uh, Ub)) == NSG_OK) {
            for (int off = 0; off < 2 && rc == NSG_OK; ++off) {
                hipLaunchKernelGGL(products_kernel, dim3(GB), dim3(TB), 0, les->s, (const double*)Ub, Pb, nl, off);
                if (!hip(hipGetLastError(), "products")) break;
                rc = nsgi_r2c(les, Pb, off ? Qo : Qd);
            }
        }
        // tau = F - Q, back to real space, deviatoric part
        double* Td = Rl + 6 * nl;
        double* To = Rl + 9 * nl;
        if (rc == NSG_OK) {
            hipLaunchKernelGGL(sub_kernel, dim3(GB), dim3(TB), 0, les->s, (const double2*)Fd, (const double2*)Qd, Qd, 3 * pll);
            hipLaunchKernelGGL(sub_kernel, dim3(GB), dim3(TB), 0, les->s, (const double2*)Fo, (const double2*)Qo, Qo, 3 * pll);
            hip(hipGetLastError(), "sub");
        }
        if (rc == NSG_OK) rc = nsgi_c2r(les, Qd, Td);
        if (rc == NSG_OK) rc = nsgi_c2r(les, Qo, To);
        if (rc == NSG_OK) {
            float* td = (float*)Rl;  // gradient scratch reused
            hipLaunchKernelGGL(tau_out_kernel, dim3(GB), dim3(TB), 0, les->s, (const double*)Td, (const double*)To, td, nl);
            if (hip(hipGetLastError(), "tau") && hip(hipMemcpyAsync(tau_out, td, sizeof(float) * 6 * nl, hipMemcpyDeviceToHost, les->s), "copy"))
                hip(hipStreamSynchronize(les->s), "sync");
        }
    }
    if (rc == NSG_OK) hip(hipStreamSynchronize(les->s), "sync");
    cleanup();
    return rc;
}
