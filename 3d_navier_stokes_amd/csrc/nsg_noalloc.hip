// nsg_noalloc.hip -- nsg_measures_noalloc: ComputeSystemMeasurables (Solver/src/solver.c:1142-1276)
// for callers that cannot allocate (code under HIP graph capture, tight serving loops, harnesses
// that forbid hipMalloc).  It runs the kernels of nsg_measures -- measures_kernel writes 7 partial
// sums per spectral shell -- but keeps those partials in a file-scope device array instead of the
// context's buffer, reduces them on the device in the host path's order, and writes totals and
// spectra to the caller's device buffer.  Nothing is allocated and nothing waits on the host.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>

#include "../../include/nsg.h"
#include "nsg_internal.hpp"

// measures_kernel (nsg_kernels.hpp) is compiled and registered by nsg_api.hip; it is launched
// from here through its declaration
namespace nsg {
__global__ void measures_kernel(const double2* __restrict__ uh0, double* __restrict__ part0, int N, int kmax, int NK,
                                int NZP, const DevScalars* __restrict__ sc, int visc_pow, int quirk, long long sbox);
}
using namespace nsg;

namespace {

// Partial sums: 7 doubles per shell, n_spect = (int)sqrt(3 (N/2)^2) + 1 shells (887 at N = 1024,
// 6209 doubles).  64 KiB per device, held until the process ends.
constexpr int kScratchDoubles = 8192;
constexpr int kMaxDevices = 64;

std::once_flag g_once[kMaxDevices];
double* g_scratch[kMaxDevices];  // device address of `scratch` on each device
hipError_t g_scratch_err[kMaxDevices];

// totals (in nsg_totals order) and spectra from the per-shell partials, summed in shell order as
// the host path of nsg_measures does (the same roundings)
__global__ void measures_finalize_kernel(const double* __restrict__ part, int ns, int n_spect, double nf, double cf,
                                         double* __restrict__ out) {
    if (threadIdx.x == 0) {
        double t[5] = {0, 0, 0, 0, 0};
        for (int q = 0; q < ns; ++q)
            for (int j = 0; j < 5; ++j) t[j] += part[q * 7 + j];
        out[0] = t[0] * cf * nf;
        out[1] = t[1] * cf * nf;
        out[2] = t[2] * cf * nf;
        out[3] = t[3] * cf * nf;
        out[4] = t[4] * 2.0 * cf * nf;
    }
    for (int q = threadIdx.x; q < n_spect; q += blockDim.x) {
        out[5 + q] = q < ns ? cf * nf * part[q * 7 + 5] : 0.0;
        out[5 + n_spect + q] = q < ns ? cf * nf * part[q * 7 + 6] : 0.0;
    }
}

}  // namespace

// the per-device scratch: in host code `scratch` names a host-side stand-in, never its memory; the
// device address comes from hipGetSymbolAddress, once per device
static __device__ __align__(256) double scratch[kScratchDoubles];

// called by nsg_create with `device` current
void nsgi_noalloc_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return;
    std::call_once(g_once[device], [device] {
        void* p = nullptr;
        g_scratch_err[device] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        g_scratch[device] = g_scratch_err[device] == hipSuccess ? (double*)p : nullptr;
    });
}

extern "C" int nsg_measures_noalloc(nsg_ctx* c, double* out, int n_spect, int quirk) {
    if (!c || !out) return NSG_ERR_INVALID;
    const int dev = c->p.device;
    if (dev < 0 || dev >= kMaxDevices) return fail(c, NSG_ERR_INVALID, "device %d outside 0..%d", dev, kMaxDevices - 1);
    if (n_spect < c->nspect) return fail(c, NSG_ERR_INVALID, "n_spect %d < %d", n_spect, c->nspect);
    const int ns = c->nspect;
    if (ns * 7 > kScratchDoubles)
        return fail(c, NSG_ERR_INVALID, "%d spectral shells need %d partial sums; the static scratch holds %d", ns,
                    ns * 7, kScratchDoubles);
    HIPCHK(c, hipSetDevice(dev));
    double* part = g_scratch[dev];
    if (!part)
        return fail(c, NSG_ERR_DEVICE, "no device address for the measures scratch on device %d: %s", dev,
                    hipGetErrorString(g_scratch_err[dev]));
    hipLaunchKernelGGL(measures_kernel, dim3(ns, 1), dim3(1024), 0, c->s, c->uh, part, c->N, c->kmax, c->NK, c->NZP,
                       (const DevScalars*)c->sc, c->p.visc_pow, quirk ? 1 : 0, c->sbox);
    LAUNCHCHK(c);
    // the normalisations of nsg_measures, formed on the host as there
    const double N3 = (double)c->N * c->N * c->N;
    const double nf = 0.5 / (N3 * N3), cf = 8.0 * M_PI * M_PI * M_PI;
    hipLaunchKernelGGL(measures_finalize_kernel, dim3(1), dim3(256), 0, c->s, (const double*)part, ns, n_spect, nf, cf,
                       out);
    LAUNCHCHK(c);
    return NSG_OK;
}
