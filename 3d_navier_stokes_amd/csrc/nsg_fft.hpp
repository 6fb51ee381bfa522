// nsg_fft.hpp -- register/LDS Stockham FFT engine for CDNA4 (gfx950).
//
// One transform of length L is carried by TPF = L/EPT threads, each holding EPT complex values in
// registers.  Stage s of radix R (R | EPT) performs EPT/R butterflies per thread; between stages
// the data goes through LDS.  The register pattern on entry and on exit is the same:
// thread t holds positions t + TPF*m, m = 0..EPT-1, which is what lets the passes load and store
// global memory directly (consecutive lanes take consecutive columns of one row).
//
// Radix-R stage (p = product of the radices already applied):
//   butterfly b reads x[b + q*L/R], twiddles by W_{pR}^{q (b mod p)}, DFT_R, and writes
//   y[(b - b mod p) R + b mod p + q p]                                  (Stockham autosort)
#pragma once
#include <hip/hip_runtime.h>

namespace nsg {

template <typename T> struct cpx;
template <> struct cpx<double> { using type = double2; };
template <> struct cpx<float> { using type = float2; };
template <typename T> using cpx_t = typename cpx<T>::type;

template <typename C> __device__ __forceinline__ C cmk(decltype(C::x) a, decltype(C::x) b) { C r; r.x = a; r.y = b; return r; }
template <typename C> __device__ __forceinline__ C cadd(C a, C b) { return cmk<C>(a.x + b.x, a.y + b.y); }
template <typename C> __device__ __forceinline__ C csub(C a, C b) { return cmk<C>(a.x - b.x, a.y - b.y); }
template <typename C> __device__ __forceinline__ C cmul(C a, C b) { return cmk<C>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
template <typename C> __device__ __forceinline__ C cconj(C a) { return cmk<C>(a.x, -a.y); }
// multiply by (DIR * i): DIR = -1 forward (e^{-}), +1 inverse (e^{+})
template <int DIR, typename C> __device__ __forceinline__ C cmul_i(C a) {
    if constexpr (DIR < 0) return cmk<C>(a.y, -a.x);  // -i a
    else return cmk<C>(-a.y, a.x);                     // +i a
}

// ------------------------------------------------------------------------------------------
// compile-time stage schedule
// ------------------------------------------------------------------------------------------
constexpr int cgcd(int a, int b) { return b == 0 ? a : cgcd(b, a % b); }
constexpr int pick_radix(int rem, int ept) {
    int g = cgcd(rem, ept);
    if (g % 8 == 0) return 8;
    if (g % 4 == 0) return 4;
    if (g % 2 == 0) return 2;
    if (g % 3 == 0) return 3;
    return 0;  // unsupported
}
constexpr int stage_radix(int L, int ept, int s) {
    int rem = L;
    for (int i = 0; i < s; ++i) { int r = pick_radix(rem, ept); if (r == 0) return 0; rem /= r; }
    return rem == 1 ? 1 : pick_radix(rem, ept);
}
constexpr int stage_p(int L, int ept, int s) {  // product of radices before stage s
    int p = 1;
    for (int i = 0; i < s; ++i) p *= stage_radix(L, ept, i);
    return p;
}
constexpr int num_stages(int L, int ept) {
    int rem = L, n = 0;
    while (rem > 1) { int r = pick_radix(rem, ept); if (r == 0) return -1; rem /= r; ++n; }
    return n;
}
constexpr bool schedule_ok(int L, int ept) { return L % ept == 0 && num_stages(L, ept) >= 0; }

// ------------------------------------------------------------------------------------------
// small DFTs on registers, in place: x[0..R-1] at stride S inside v[]
// ------------------------------------------------------------------------------------------
template <int DIR, typename C>
__device__ __forceinline__ void dft2(C& a, C& b) { C t = a; a = cadd(t, b); b = csub(t, b); }

template <int DIR, typename C>
__device__ __forceinline__ void dft4(C& x0, C& x1, C& x2, C& x3) {
    C a = cadd(x0, x2), b = csub(x0, x2), c = cadd(x1, x3), d = cmul_i<DIR>(csub(x1, x3));
    x0 = cadd(a, c); x2 = csub(a, c);
    x1 = cadd(b, d); x3 = csub(b, d);
}

template <int DIR, typename C>
__device__ __forceinline__ void dft3(C& x0, C& x1, C& x2) {
    using R = decltype(C::x);
    const R c1 = R(-0.5);
    const R s1 = R(DIR) * R(0.86602540378443864676372317075294);  // DIR * sin(2pi/3)
    C t1 = cadd(x1, x2), t2 = csub(x1, x2);
    C m1 = cmk<C>(x0.x + c1 * t1.x, x0.y + c1 * t1.y);
    C m2 = cmk<C>(-s1 * t2.y, s1 * t2.x);  // i * s1 * t2
    x0 = cadd(x0, t1);
    x1 = cadd(m1, m2);
    x2 = csub(m1, m2);
}

template <int DIR, typename C>
__device__ __forceinline__ void dft8(C& x0, C& x1, C& x2, C& x3, C& x4, C& x5, C& x6, C& x7) {
    using R = decltype(C::x);
    const R h = R(0.70710678118654752440084436210485);
    dft4<DIR>(x0, x2, x4, x6);  // E
    dft4<DIR>(x1, x3, x5, x7);  // O
    // O[k] *= W8^{k DIR}
    {
        C t = x3;  // k=1: (1 + DIR i) h  ... W8^{-1} forward = (1 - i)/sqrt2
        if constexpr (DIR < 0) x3 = cmk<C>(h * (t.x + t.y), h * (t.y - t.x));
        else x3 = cmk<C>(h * (t.x - t.y), h * (t.y + t.x));
    }
    x5 = cmul_i<DIR>(x5);  // k=2
    {
        C t = x7;  // k=3: W8^{3} forward = (-1 - i)/sqrt2
        if constexpr (DIR < 0) x7 = cmk<C>(h * (-t.x + t.y), h * (-t.y - t.x));
        else x7 = cmk<C>(h * (-t.x - t.y), h * (-t.y + t.x));
    }
    C e0 = x0, e1 = x2, e2 = x4, e3 = x6, o0 = x1, o1 = x3, o2 = x5, o3 = x7;
    x0 = cadd(e0, o0); x4 = csub(e0, o0);
    x1 = cadd(e1, o1); x5 = csub(e1, o1);
    x2 = cadd(e2, o2); x6 = csub(e2, o2);
    x3 = cadd(e3, o3); x7 = csub(e3, o3);
}

// ------------------------------------------------------------------------------------------
// twiddle fetch: tw[e] = exp(-2 pi i e / L) table (L entries), conjugated for the inverse
// ------------------------------------------------------------------------------------------
template <int DIR, typename C>
__device__ __forceinline__ C twiddle(const C* __restrict__ tw, int e) {
    C w = tw[e];
    if constexpr (DIR > 0) w.y = -w.y;
    return w;
}

// One radix-R stage on the registers of one thread.  v holds inputs at element m = beta + B*q.
template <int L, int EPT, int TPF, int S, int DIR, typename C>
__device__ __forceinline__ void stage_compute(C (&v)[EPT], int t, const C* __restrict__ tw) {
    constexpr int R = stage_radix(L, EPT, S);
    constexpr int P = stage_p(L, EPT, S);
    constexpr int B = EPT / R;
#pragma unroll
    for (int beta = 0; beta < B; ++beta) {
        const int b = t + TPF * beta;
        if constexpr (P > 1) {
            const int k = b % P;
#pragma unroll
            for (int q = 1; q < R; ++q) v[beta + B * q] = cmul(v[beta + B * q], twiddle<DIR>(tw, q * k * (L / (P * R))));
        }
        if constexpr (R == 2) dft2<DIR>(v[beta], v[beta + B]);
        else if constexpr (R == 3) dft3<DIR>(v[beta], v[beta + B], v[beta + 2 * B]);
        else if constexpr (R == 4) dft4<DIR>(v[beta], v[beta + B], v[beta + 2 * B], v[beta + 3 * B]);
        else if constexpr (R == 8)
            dft8<DIR>(v[beta], v[beta + B], v[beta + 2 * B], v[beta + 3 * B], v[beta + 4 * B], v[beta + 5 * B],
                      v[beta + 6 * B], v[beta + 7 * B]);
    }
}

// LDS slot of position pos in the exchange that follows a stage (P, R).  With PAD, when the
// stage writes runs of P < 16 contiguous positions spaced P*R apart, P pad slots follow every
// P*R positions so that neighbouring lanes fall in different LDS banks.
template <int P, int R, bool PAD> __device__ __forceinline__ int lds_slot(int pos) {
    if constexpr (PAD && P < 16) return pos + P * (pos / (P * R));
    else return pos;
}
// LDS elements one transform needs (padded exchanges need up to L + L/R slots)
constexpr int fft_lds_len(int L, int ept, bool pad) {
    int n = L;
    if (!pad) return n;
    for (int s = 0; s + 1 < num_stages(L, ept); ++s) {
        int P = stage_p(L, ept, s), R = stage_radix(L, ept, s);
        if (P < 16) { int m = L + P * ((L - 1) / (P * R)) + 1; if (m > n) n = m; }
    }
    return n;
}

// write stage outputs to LDS at the Stockham output positions; lds address = base + slot*PS
template <int L, int EPT, int TPF, int S, bool PAD, typename C>
__device__ __forceinline__ void stage_scatter(const C (&v)[EPT], int t, C* lds, int base, int PS) {
    constexpr int R = stage_radix(L, EPT, S);
    constexpr int P = stage_p(L, EPT, S);
    constexpr int B = EPT / R;
#pragma unroll
    for (int beta = 0; beta < B; ++beta) {
        const int b = t + TPF * beta;
        const int k = b % P;
        const int o = (b - k) * R + k;
#pragma unroll
        for (int q = 0; q < R; ++q) lds[base + lds_slot<P, R, PAD>(o + q * P) * PS] = v[beta + B * q];
    }
}

template <int L, int EPT, int TPF, int S, bool PAD, typename C>
__device__ __forceinline__ void gather_pattern(C (&v)[EPT], int t, const C* lds, int base, int PS) {
    constexpr int R = stage_radix(L, EPT, S);
    constexpr int P = stage_p(L, EPT, S);
#pragma unroll
    for (int m = 0; m < EPT; ++m) v[m] = lds[base + lds_slot<P, R, PAD>(t + TPF * m) * PS];
}

// LDS exchange barrier.  SYNC_BLOCK: the transform spans several wavefronts (s_barrier);
// SYNC_WAVE: all TPF threads of a transform sit in one wavefront, which executes its LDS
// instructions in order, so only a wavefront-scope fence is needed and waves run independently;
// SYNC_LDS: a block barrier that waits for the wavefront's own LDS accesses only, so global reads
// and global->LDS copies into other LDS buffers stay in flight across the transform (the caller
// keeps every LDS access of the transform away from those buffers).
enum SyncKind { SYNC_BLOCK = 0, SYNC_WAVE = 1, SYNC_LDS = 2 };
template <int SYNC> __device__ __forceinline__ void lds_sync() {
    if constexpr (SYNC == SYNC_BLOCK) {
        __syncthreads();
    } else if constexpr (SYNC == SYNC_LDS) {
        asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
    } else {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

template <int L, int EPT, int TPF, int S, int NS, int DIR, int SYNC, bool PAD, typename C>
__device__ __forceinline__ void fft_stages(C (&v)[EPT], int t, C* lds, int base, int PS, const C* __restrict__ tw) {
    if constexpr (S < NS) {
        stage_compute<L, EPT, TPF, S, DIR>(v, t, tw);
        if constexpr (S + 1 < NS) {
            stage_scatter<L, EPT, TPF, S, PAD>(v, t, lds, base, PS);
            lds_sync<SYNC>();
            gather_pattern<L, EPT, TPF, S, PAD>(v, t, lds, base, PS);
            lds_sync<SYNC>();
        }
        fft_stages<L, EPT, TPF, S + 1, NS, DIR, SYNC, PAD>(v, t, lds, base, PS, tw);
    }
}

// Full length-L transform of the values held in pattern t + TPF*m; result in the same pattern.
// With SYNC_BLOCK every thread of the block must call it (it contains barriers).
template <int L, int EPT, int DIR, int SYNC = SYNC_BLOCK, bool PAD = false, typename C>
__device__ __forceinline__ void fft(C (&v)[EPT], int t, C* lds, int base, int PS, const C* __restrict__ tw) {
    static_assert(schedule_ok(L, EPT), "unsupported FFT length / elements-per-thread");
    constexpr int TPF = L / EPT;
    constexpr int NS = num_stages(L, EPT);
    static_assert(SYNC != SYNC_WAVE || (TPF <= 64 && 64 % TPF == 0), "wave-synchronous transform must fit a wavefront");
    fft_stages<L, EPT, TPF, 0, NS, DIR, SYNC, PAD>(v, t, lds, base, PS, tw);
}

// ------------------------------------------------------------------------------------------
// Register-twiddle variant for transforms whose threads keep the same t across many transforms
// (the persistent z pass): the base twiddle W_{PR}^{b mod P} of every butterfly of stages >= 1 is
// loaded once (fft_load_base_twiddles) and its powers q = 2..R-1 are formed by multiplication.
// ------------------------------------------------------------------------------------------
constexpr int num_base_twiddles(int L, int ept) {
    int n = 0;
    for (int s = 1; s < num_stages(L, ept); ++s) n += ept / stage_radix(L, ept, s);
    return n;
}
template <int L, int EPT, int S> constexpr int base_twiddle_offset() {
    int n = 0;
    for (int s = 1; s < S; ++s) n += EPT / stage_radix(L, EPT, s);
    return n;
}

template <int L, int EPT, int TPF, int S, int NTW, typename C>
__device__ __forceinline__ void load_base_stage(C (&w)[NTW], int t, const C* __restrict__ tw) {
    constexpr int NS = num_stages(L, EPT);
    if constexpr (S < NS) {
        constexpr int R = stage_radix(L, EPT, S);
        constexpr int P = stage_p(L, EPT, S);
        constexpr int B = EPT / R;
        constexpr int off = base_twiddle_offset<L, EPT, S>();
#pragma unroll
        for (int beta = 0; beta < B; ++beta) w[off + beta] = tw[((t + TPF * beta) % P) * (L / (P * R))];
        load_base_stage<L, EPT, TPF, S + 1, NTW>(w, t, tw);
    }
}
// w[] receives exp(-2 pi i (b mod P)/(P R)) for every butterfly of stages 1..NS-1 (forward sign)
template <int L, int EPT, typename C>
__device__ __forceinline__ void fft_load_base_twiddles(C (&w)[num_base_twiddles(L, EPT) > 0 ? num_base_twiddles(L, EPT) : 1],
                                                       int t, const C* __restrict__ tw) {
    constexpr int TPF = L / EPT;
    load_base_stage<L, EPT, TPF, 1, (num_base_twiddles(L, EPT) > 0 ? num_base_twiddles(L, EPT) : 1)>(w, t, tw);
}

template <int L, int EPT, int TPF, int S, int DIR, int NTW, typename C>
__device__ __forceinline__ void stage_compute_reg(C (&v)[EPT], const C (&w)[NTW]) {
    constexpr int R = stage_radix(L, EPT, S);
    constexpr int B = EPT / R;
    constexpr int off = (S > 0) ? base_twiddle_offset<L, EPT, S>() : 0;
#pragma unroll
    for (int beta = 0; beta < B; ++beta) {
        if constexpr (S > 0) {
            C w1 = w[off + beta];
            if constexpr (DIR > 0) w1.y = -w1.y;
            C wq = w1;
#pragma unroll
            for (int q = 1; q < R; ++q) {
                v[beta + B * q] = cmul(v[beta + B * q], wq);
                if (q + 1 < R) wq = cmul(wq, w1);
            }
        }
        if constexpr (R == 2) dft2<DIR>(v[beta], v[beta + B]);
        else if constexpr (R == 3) dft3<DIR>(v[beta], v[beta + B], v[beta + 2 * B]);
        else if constexpr (R == 4) dft4<DIR>(v[beta], v[beta + B], v[beta + 2 * B], v[beta + 3 * B]);
        else if constexpr (R == 8)
            dft8<DIR>(v[beta], v[beta + B], v[beta + 2 * B], v[beta + 3 * B], v[beta + 4 * B], v[beta + 5 * B],
                      v[beta + 6 * B], v[beta + 7 * B]);
    }
}

template <int L, int EPT, int TPF, int S, int NS, int DIR, int SYNC, bool PAD, int NTW, typename C>
__device__ __forceinline__ void fft_stages_reg(C (&v)[EPT], int t, C* lds, int base, int PS, const C (&w)[NTW]) {
    if constexpr (S < NS) {
        stage_compute_reg<L, EPT, TPF, S, DIR, NTW>(v, w);
        if constexpr (S + 1 < NS) {
            stage_scatter<L, EPT, TPF, S, PAD>(v, t, lds, base, PS);
            lds_sync<SYNC>();
            gather_pattern<L, EPT, TPF, S, PAD>(v, t, lds, base, PS);
            lds_sync<SYNC>();
        }
        fft_stages_reg<L, EPT, TPF, S + 1, NS, DIR, SYNC, PAD, NTW>(v, t, lds, base, PS, w);
    }
}

template <int L, int EPT, int DIR, int SYNC, bool PAD, int NTW, typename C>
__device__ __forceinline__ void fft_reg(C (&v)[EPT], int t, C* lds, int base, int PS, const C (&w)[NTW]) {
    static_assert(schedule_ok(L, EPT), "unsupported FFT length / elements-per-thread");
    constexpr int TPF = L / EPT;
    fft_stages_reg<L, EPT, TPF, 0, num_stages(L, EPT), DIR, SYNC, PAD, NTW>(v, t, lds, base, PS, w);
}

// ------------------------------------------------------------------------------------------
// Full-table variant: every twiddle W_{PR}^{q (b mod P)} (q = 1..R-1) of every butterfly of
// stages >= 1 is held in registers (fft_load_twiddle_table, once per kernel), so a transform issues
// no twiddle arithmetic besides the products themselves.  Costs sum_s (R_s - 1) EPT / R_s complex
// registers per thread.
// ------------------------------------------------------------------------------------------
constexpr int num_table_twiddles(int L, int ept) {
    int n = 0;
    for (int s = 1; s < num_stages(L, ept); ++s) n += (stage_radix(L, ept, s) - 1) * (ept / stage_radix(L, ept, s));
    return n;
}
constexpr int ntab(int L, int ept) { return num_table_twiddles(L, ept) > 0 ? num_table_twiddles(L, ept) : 1; }
template <int L, int EPT, int S> constexpr int table_twiddle_offset() {
    int n = 0;
    for (int s = 1; s < S; ++s) n += (stage_radix(L, EPT, s) - 1) * (EPT / stage_radix(L, EPT, s));
    return n;
}

template <int L, int EPT, int TPF, int S, typename C>
__device__ __forceinline__ void load_table_stage(C (&w)[ntab(L, EPT)], int t, const C* __restrict__ tw) {
    if constexpr (S < num_stages(L, EPT)) {
        constexpr int R = stage_radix(L, EPT, S);
        constexpr int P = stage_p(L, EPT, S);
        constexpr int B = EPT / R;
        constexpr int off = table_twiddle_offset<L, EPT, S>();
#pragma unroll
        for (int beta = 0; beta < B; ++beta) {
            const int k = (t + TPF * beta) % P;
#pragma unroll
            for (int q = 1; q < R; ++q) w[off + beta * (R - 1) + q - 1] = tw[q * k * (L / (P * R))];
        }
        load_table_stage<L, EPT, TPF, S + 1>(w, t, tw);
    }
}
// w[] receives exp(-2 pi i q (b mod P)/(P R)) for q = 1..R-1 and every butterfly b of stages
// 1..NS-1 (forward sign; the inverse conjugates on use); tw is the length-L table exp(-2 pi i e/L)
template <int L, int EPT, typename C>
__device__ __forceinline__ void fft_load_twiddle_table(C (&w)[ntab(L, EPT)], int t, const C* __restrict__ tw) {
    load_table_stage<L, EPT, L / EPT, 1>(w, t, tw);
}

template <int L, int EPT, int S, int DIR, typename C>
__device__ __forceinline__ void stage_compute_tab(C (&v)[EPT], const C (&w)[ntab(L, EPT)]) {
    constexpr int R = stage_radix(L, EPT, S);
    constexpr int B = EPT / R;
#pragma unroll
    for (int beta = 0; beta < B; ++beta) {
        if constexpr (S > 0) {
            constexpr int off = table_twiddle_offset<L, EPT, S>();
#pragma unroll
            for (int q = 1; q < R; ++q) {
                C wq = w[off + beta * (R - 1) + q - 1];
                if constexpr (DIR > 0) wq.y = -wq.y;
                v[beta + B * q] = cmul(v[beta + B * q], wq);
            }
        }
        if constexpr (R == 2) dft2<DIR>(v[beta], v[beta + B]);
        else if constexpr (R == 3) dft3<DIR>(v[beta], v[beta + B], v[beta + 2 * B]);
        else if constexpr (R == 4) dft4<DIR>(v[beta], v[beta + B], v[beta + 2 * B], v[beta + 3 * B]);
        else if constexpr (R == 8)
            dft8<DIR>(v[beta], v[beta + B], v[beta + 2 * B], v[beta + 3 * B], v[beta + 4 * B], v[beta + 5 * B],
                      v[beta + 6 * B], v[beta + 7 * B]);
    }
}

template <int L, int EPT, int TPF, int S, int NS, int DIR, int SYNC, bool PAD, typename C>
__device__ __forceinline__ void fft_stages_tab(C (&v)[EPT], int t, C* lds, int base, int PS, const C (&w)[ntab(L, EPT)]) {
    if constexpr (S < NS) {
        stage_compute_tab<L, EPT, S, DIR>(v, w);
        if constexpr (S + 1 < NS) {
            stage_scatter<L, EPT, TPF, S, PAD>(v, t, lds, base, PS);
            lds_sync<SYNC>();
            gather_pattern<L, EPT, TPF, S, PAD>(v, t, lds, base, PS);
            lds_sync<SYNC>();
        }
        fft_stages_tab<L, EPT, TPF, S + 1, NS, DIR, SYNC, PAD>(v, t, lds, base, PS, w);
    }
}

template <int L, int EPT, int DIR, int SYNC, bool PAD, typename C>
__device__ __forceinline__ void fft_tab(C (&v)[EPT], int t, C* lds, int base, int PS, const C (&w)[ntab(L, EPT)]) {
    static_assert(schedule_ok(L, EPT), "unsupported FFT length / elements-per-thread");
    constexpr int TPF = L / EPT;
    fft_stages_tab<L, EPT, TPF, 0, num_stages(L, EPT), DIR, SYNC, PAD>(v, t, lds, base, PS, w);
}

}  // namespace nsg
