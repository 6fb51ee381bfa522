// CPU emulation of the persistent z passes (test infrastructure): compiles the z-pass kernels from
// 3d_navier_stokes_amd/csrc/nsg_kernels.hpp with g++ against tests/emu/hip/hip_runtime.h and runs
// it with one coroutine per lane, one wavefront at a time (see that header for the lockstep rules).
// tests/test_zpass_emulation.py loads this as a shared library and checks the result line by line
// against numpy's irfft / rfft, for every (precision, N, elements per thread) the library builds
// and for the FP64 N = 16, 8-elements-per-thread instance that the library no longer builds.
#include <cmath>
#include <vector>

#include "../3d_navier_stokes_amd/csrc/nsg_kernels.hpp"

using namespace nsg;

namespace {

// a grid of `blocks` blocks of `threads` threads: each block's wavefronts run one after another
template <typename F> void launch(F kern, const PZArgs& a, unsigned blocks, unsigned threads) {
    static emu::Wave wave;
    for (unsigned b = 0; b < blocks; ++b) {
        emu::t_bid = {b, 0, 0};
        emu::t_gdim = {blocks, 1, 1};
        emu::t_bdim = {threads, 1, 1};
        for (unsigned w0 = 0; w0 < threads; w0 += 64)
            emu::run_wave(wave, (int)w0, (int)std::min(64u, threads - w0), [&] { kern(a); });
    }
}

template <typename T> std::vector<cpx_t<T>> table(int L) {
    std::vector<cpx_t<T>> t(L);
    for (int e = 0; e < L; ++e) {
        const double a = 2.0 * kPi * (double)e / (double)L;
        t[e].x = (T)std::cos(a);
        t[e].y = (T)-std::sin(a);
    }
    return t;
}

// ver 1: passz_rhs_kernel; 2: passz_rhs2_kernel (base twiddles); 3: passz_rhs2_kernel (twiddle table);
// 4: passz_full_kernel (base twiddles, P_z of two lines per transform where the line groups pair up);
// 5: passz_full_kernel (twiddle table, P_z alone); 6: passz_full_kernel (twiddle table, P_z paired).
// eptz is the elements per thread of that kernel's transform (N/2 points for 1-3, N points for 4-6).
template <typename T, int N, int EPTZ>
int run(int ver, int nbox, unsigned blocks, const void* in, void* out, double scale, double* maxout) {
    constexpr int WPB = 4;
    auto twM = table<T>(N / 2);
    auto twN = table<T>(N);
    std::vector<DevScalars> sc(nbox);
    for (auto& s : sc) std::memset(&s, 0, sizeof s);
    PZArgs a;
    std::memset(&a, 0, sizeof a);
    a.in = in;
    a.out = out;
    a.twM = twM.data();
    a.twN = twN.data();
    a.sc = sc.data();
    a.scale = scale;
    a.ibox = 6LL * N * N * Ext<N>::NZP;
    a.obox = 3LL * N * N * Ext<N>::NZP;
    a.nbox = nbox;
    a.kmax = Ext<N>::kmax;
    a.NZP = Ext<N>::NZP;
    a.track_max = 1;
    constexpr bool half_ok = (N / 2) % EPTZ == 0 && (N / 2) / EPTZ <= 64 && 64 % ((N / 2) / EPTZ) == 0 &&
                             ((N / 2) & (N / 2 - 1)) == 0 && schedule_ok(N / 2, EPTZ);
    constexpr bool full_ok = N % EPTZ == 0 && N / EPTZ <= 64 && 64 % (N / EPTZ) == 0 && schedule_ok(N, EPTZ);
    if (ver <= 3) {
        if constexpr (half_ok) {
            if (ver == 1) launch(passz_rhs_kernel<T, N, EPTZ, WPB, 1>, a, blocks, 64 * WPB);
            else if (ver == 2) launch(passz_rhs2_kernel<T, N, EPTZ, WPB, 1, false>, a, blocks, 64 * WPB);
            else launch(passz_rhs2_kernel<T, N, EPTZ, WPB, 1, true>, a, blocks, 64 * WPB);
        } else {
            return -1;
        }
    } else if (ver <= 6) {
        if constexpr (full_ok) {
            if (ver == 4) launch(passz_full_kernel<T, N, EPTZ, WPB, 1, false, true>, a, blocks, 64 * WPB);
            else if (ver == 5) launch(passz_full_kernel<T, N, EPTZ, WPB, 1, true, false>, a, blocks, 64 * WPB);
            else launch(passz_full_kernel<T, N, EPTZ, WPB, 1, true, true>, a, blocks, 64 * WPB);
        } else {
            return -1;
        }
    } else {
        return -1;
    }
    for (int b = 0; b < nbox; ++b) maxout[b] = __longlong_as_double((long long)sc[b].maxbits);
    return 0;
}

template <typename T, int N> int by_ept(int ver, int eptz, int nbox, unsigned blocks, const void* in, void* out,
                                        double scale, double* mx) {
    switch (eptz) {
        case 2: return run<T, N, 2>(ver, nbox, blocks, in, out, scale, mx);
        case 4: if constexpr (N % 4 == 0) return run<T, N, 4>(ver, nbox, blocks, in, out, scale, mx); break;
        case 8: if constexpr (N % 8 == 0) return run<T, N, 8>(ver, nbox, blocks, in, out, scale, mx); break;
        case 16: if constexpr (N >= 16 && N % 16 == 0) return run<T, N, 16>(ver, nbox, blocks, in, out, scale, mx); break;
        case 3: if constexpr (N % 3 == 0) return run<T, N, 3>(ver, nbox, blocks, in, out, scale, mx); break;
        case 6: if constexpr (N % 6 == 0) return run<T, N, 6>(ver, nbox, blocks, in, out, scale, mx); break;
        case 12: if constexpr (N % 12 == 0) return run<T, N, 12>(ver, nbox, blocks, in, out, scale, mx); break;
        default: break;
    }
    return -1;
}

}  // namespace

extern "C" {
// the Ext<N> extents the emulated kernel uses: kmax, NZP
void emu_extents(int N, int* kmax, int* nzp) {
    *kmax = N / 3;
    *nzp = ((N / 3 + 1 + 15) / 16) * 16;
}
// z pass of nbox boxes: in [nbox][6][N][N][NZP], out [nbox][3][N][N][NZP] (complex, fp64 or fp32)
int emu_passz_rhs(int N, int fp32, int eptz, int ver, int nbox, int blocks, const void* in, void* out, double scale,
                  double* max_per_box) {
    auto go = [&](auto tag) -> int {
        using T = decltype(tag);
        switch (N) {
            case 8: return by_ept<T, 8>(ver, eptz, nbox, blocks, in, out, scale, max_per_box);
            case 16: return by_ept<T, 16>(ver, eptz, nbox, blocks, in, out, scale, max_per_box);
            case 32: return by_ept<T, 32>(ver, eptz, nbox, blocks, in, out, scale, max_per_box);
            case 64: return by_ept<T, 64>(ver, eptz, nbox, blocks, in, out, scale, max_per_box);
            case 128: return by_ept<T, 128>(ver, eptz, nbox, blocks, in, out, scale, max_per_box);
            case 12: return by_ept<T, 12>(ver, eptz, nbox, blocks, in, out, scale, max_per_box);
            case 24: return by_ept<T, 24>(ver, eptz, nbox, blocks, in, out, scale, max_per_box);
            case 48: return by_ept<T, 48>(ver, eptz, nbox, blocks, in, out, scale, max_per_box);
            case 96: return by_ept<T, 96>(ver, eptz, nbox, blocks, in, out, scale, max_per_box);
            default: return -1;
        }
    };
    return fp32 ? go(0.0f) : go(0.0);
}
}
