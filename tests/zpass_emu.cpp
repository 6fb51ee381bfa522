// CPU emulation of the persistent z pass (test infrastructure): compiles passz_rhs_kernel from
// 3d_navier_stokes_amd/csrc/nsg_kernels.hpp with g++ against tests/emu/hip/hip_runtime.h and runs
// it with one host thread per lane, one block at a time (see that header for the lockstep rules).
// tests/test_zpass_emulation.py loads this as a shared library and checks the result line by line
// against numpy's irfft / rfft, for every (precision, N, elements per thread) the library builds
// and for the FP64 N = 16, 8-elements-per-thread instance that the library no longer builds.
#include <cmath>
#include <memory>
#include <thread>
#include <vector>

#include "../3d_navier_stokes_amd/csrc/nsg_kernels.hpp"

using namespace nsg;

namespace {

template <typename F> void launch(F kern, const PZArgs& a, unsigned blocks, unsigned threads) {
    for (unsigned b = 0; b < blocks; ++b) {
        const unsigned nw = (threads + 63) / 64;
        std::vector<emu::WaveCtx> waves(nw);
        std::vector<std::unique_ptr<std::barrier<>>> wbar;
        for (unsigned w = 0; w < nw; ++w) {
            const unsigned lanes = std::min(64u, threads - 64 * w);
            wbar.emplace_back(new std::barrier<>((std::ptrdiff_t)lanes));
            waves[w].bar = wbar.back().get();
        }
        std::barrier<> blockbar((std::ptrdiff_t)threads);
        std::vector<std::thread> th;
        th.reserve(threads);
        for (unsigned t = 0; t < threads; ++t)
            th.emplace_back([&, t, b] {
                emu::t_tid = {t, 0, 0};
                emu::t_bid = {b, 0, 0};
                emu::t_gdim = {blocks, 1, 1};
                emu::t_bdim = {threads, 1, 1};
                emu::t_wave = &waves[t / 64];
                emu::t_block = &blockbar;
                kern(a);
            });
        for (auto& x : th) x.join();
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

template <typename T, int N, int EPTZ>
int run(int nbox, unsigned blocks, const void* in, void* out, double scale, double* maxout) {
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
    launch(passz_rhs_kernel<T, N, EPTZ, WPB, 1>, a, blocks, 64 * WPB);
    for (int b = 0; b < nbox; ++b) maxout[b] = __longlong_as_double((long long)sc[b].maxbits);
    return 0;
}

template <typename T, int N> int by_ept(int eptz, int nbox, unsigned blocks, const void* in, void* out, double scale,
                                        double* mx) {
    constexpr int M = N / 2;
    switch (eptz) {
        case 2: if constexpr (M / 2 <= 64 && M >= 2) return run<T, N, 2>(nbox, blocks, in, out, scale, mx); break;
        case 4: if constexpr (M / 4 <= 64 && M >= 4) return run<T, N, 4>(nbox, blocks, in, out, scale, mx); break;
        case 8: if constexpr (M / 8 <= 64 && M >= 8) return run<T, N, 8>(nbox, blocks, in, out, scale, mx); break;
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
int emu_passz_rhs(int N, int fp32, int eptz, int nbox, int blocks, const void* in, void* out, double scale,
                  double* max_per_box) {
    auto go = [&](auto tag) -> int {
        using T = decltype(tag);
        switch (N) {
            case 8: return by_ept<T, 8>(eptz, nbox, blocks, in, out, scale, max_per_box);
            case 16: return by_ept<T, 16>(eptz, nbox, blocks, in, out, scale, max_per_box);
            case 32: return by_ept<T, 32>(eptz, nbox, blocks, in, out, scale, max_per_box);
            case 64: return by_ept<T, 64>(eptz, nbox, blocks, in, out, scale, max_per_box);
            case 128: return by_ept<T, 128>(eptz, nbox, blocks, in, out, scale, max_per_box);
            default: return -1;
        }
    };
    return fp32 ? go(0.0f) : go(0.0);
}
}
