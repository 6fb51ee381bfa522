// nsg_kernels.hpp -- device kernels of the pseudo-spectral time step (one periodic box per call).
//
// Spectral data live only on the 2/3-rule box |kx|,|ky|,kz <= kmax = floor(N/3) (the reference
// zeroes every mode with |k| > kmax, Solver/src/solver.c:1731-1737, so no other mode is ever
// nonzero).  Compact index of a wavenumber k in [-kmax,kmax]: k >= 0 ? k : k + NK, NK = 2 kmax + 1.
// Component-major layouts (complex elements, kz fastest, kz padded to NZP):
//   state / K_s : [3][NK][NK][NZP]     (state always FP64; K_s in the bulk precision)
//   A1 (pass 1) : [6][N ][NK][NZP]     after the inverse x transform of (u_hat, w_hat)
//   A2 (pass 2) : [6][N ][N ][NZP]     after the inverse y transform
//   A3 (pass 3) : [3][N ][N ][NZP]     r2c_z(u x w)
//   A4 (pass 4) : [3][N ][NK][NZP]     after the forward y transform
// Pass 5 performs the forward x transform and the epilogue of NonlinearRHSBatch
// (Solver/src/solver.c:689-727) fused with the RK stage bookkeeping of RK4Step (:505-608).
#pragma once
#include <type_traits>

#include "nsg_fft.hpp"

namespace nsg {

constexpr double kPi = 3.14159265358979323846;

struct DevScalars {
    double dt;                   // current time step
    double t;                    // simulation time
    unsigned long long maxbits;  // running max of |ux|+|uy|+|uz| (bit pattern of a nonneg double)
    double max_last;             // max of the last completed step (1/N^3 normalised)
    long long steps;             // completed steps
    double cfl;                  // > 0: adaptive dt
    double dx;                   // 2 pi / N
    double nu;                   // viscosity of this box
    double dperr;                // Dormand-Prince: error norm of the last try
    long long dpacc;             // Dormand-Prince: 1 if the last try was accepted
};

// ------------------------------------------------------------------------------------------
// layouts
//   state / K_s : [3][NK (ky)][NK (kx)][NZP]  -- y-major, so an x pass at fixed ky stays inside
//                 one NK*NZP plane (x-major state rows 1 MB apart thrash the TLB: 5e6 misses
//                 per launch at 512^3)
//   A1 / A4     : [F][N][NK][NZP]             -- x-major; measured faster than x-row tiles of
//                 4..32 rows or the y-major order, which slow the y passes more than they help x
// ------------------------------------------------------------------------------------------
// A1/A4 x-plane stride = NK*NZP + XPAD elements (XPAD breaks the alignment of x rows written by
// one store instruction)
constexpr int XPAD = 0;
__host__ __device__ __forceinline__ long long xplane(int NK, int NZP) { return (long long)NK * NZP + XPAD; }
// The 2/3 cut kmax = N/3, the compact extent NK = 2 kmax + 1 and the padded kz extent NZP
// (kmax + 1 rounded up to 16) are functions of N; the context sets the same values.  As
// compile-time constants they remove the loads, stores and branches of positions that lie outside
// the cut for every thread: the RHS z pass and the FP32 pass 1 use them (measured faster there;
// the other passes measured the same or slower and keep the run-time values).
template <int N> struct Ext {
    static constexpr int kmax = N / 3, NK = 2 * (N / 3) + 1, NZP = ((N / 3 + 1 + 15) / 16) * 16;
};
__device__ __forceinline__ long long sidx(int ix, int iy, int kz, int NK, int NZP) {
    return ((long long)iy * NK + ix) * NZP + kz;
}
template <int N> __device__ __forceinline__ long long xidx(int x, int iy, int kz, int NK, int NZP) {
    return (long long)x * xplane(NK, NZP) + (long long)iy * NZP + kz;
}

// Element i of an array whose base is wave-uniform, addressed by a 32-bit byte offset when O32
// (the array must then be shorter than 4 GiB from base): the compiler keeps the base in SGPRs and
// one VGPR per address (the saddr form of the global memory instructions) instead of a 64-bit
// address pair per access, which is what bounds the register budget of the x passes.
template <bool O32, typename E> __device__ __forceinline__ E& at(E* base, long long i) {
    if constexpr (O32) {
        using B = typename std::conditional<std::is_const<E>::value, const char, char>::type;
        return *(E*)((B*)base + (unsigned)((unsigned)i * (unsigned)sizeof(E)));
    } else {
        return base[i];
    }
}

// ------------------------------------------------------------------------------------------
// per-mode helpers
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ int k_of_pos(int p, int N) { return p <= N / 2 ? p : p - N; }
__device__ __forceinline__ int pos_of_k(int k, int N) { return k >= 0 ? k : k + N; }
__device__ __forceinline__ int cidx_of_k(int k, int NK) { return k >= 0 ? k : k + NK; }
__device__ __forceinline__ int k_of_cidx(int c, int kmax, int NK) { return c <= kmax ? c : c - NK; }

__device__ __forceinline__ double visc_rate(double nu, int p, long long k2) {
    double kk = (double)k2, r = kk;
    for (int q = 1; q < p; ++q) r *= kk;
    return nu * r;
}

template <typename T> __device__ __forceinline__ double2 to_d2(cpx_t<T> v) { return make_double2((double)v.x, (double)v.y); }
template <typename T> __device__ __forceinline__ cpx_t<T> from_d2(double2 v) { return cmk<cpx_t<T>>((T)v.x, (T)v.y); }

enum Scheme { RK4CN = 0, IFRK4 = 1, RK5CN = 2, DP5 = 3 };

// Dormand-Prince family: stage s input u + dt sum_{j<s} a_j K_j over the stored slopes
// (Solver/src/solver.c:236-334); the row's coefficients come from the host's tableau
constexpr int MAXSLOPES = 6;
struct StageRow {
    const void* k[MAXSLOPES];  // K_1..K_6 of the first box (bulk precision)
    double a[MAXSLOPES];       // row of the tableau (0: slope not read)
};

// ud += dt sum_j a_j K_j for j < nterm (uniform branches; fully unrolled so nothing is indexed at
// run time)
template <typename T>
__device__ __forceinline__ void add_slopes(double2 (&ud)[3], const StageRow& r, int nterm, long long boxoff,
                                           long long idx, long long plane, double dt) {
#pragma unroll
    for (int j = 0; j < MAXSLOPES; ++j) {
        if (j >= nterm || r.a[j] == 0.0) continue;
        const double cj = r.a[j] * dt;
        const cpx_t<T>* kj = (const cpx_t<T>*)r.k[j] + boxoff;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double2 kv = to_d2<T>(kj[c * plane + idx]);
            ud[c].x += cj * kv.x;
            ud[c].y += cj * kv.y;
        }
    }
}

// stage input u~_s = alpha u_hat + beta K_{s-1}
__device__ __forceinline__ void stage_in_coef(int scheme, int stage, double D, double dt, double& a, double& b) {
    if (stage <= 1) { a = 1.0; b = 0.0; return; }
    if (scheme == RK4CN) {
        a = 1.0;
        b = (stage == 4) ? dt : 0.5 * dt;
    } else {
        const double e2 = exp(-D * dt * 0.5);
        if (stage == 2) { a = e2; b = e2 * 0.5 * dt; }
        else if (stage == 3) { a = e2; b = 0.5 * dt; }
        else { a = exp(-D * dt); b = dt * e2; }
    }
}


// max that propagates NaN (a NaN state must reach the host's non-finite check)
__device__ __forceinline__ double nanmax(double a, double b) { return (a >= b || a != a) ? a : b; }

// deterministic block sum (fixed tree); red must hold blockDim.x doubles
__device__ __forceinline__ double block_sum(double v, double* red) {
    const int tid = threadIdx.x, n = blockDim.x;
    red[tid] = v;
    __syncthreads();
    int p = 1;
    while (p < n) p <<= 1;
    for (int s = p >> 1; s > 0; s >>= 1) {
        if (tid < s && tid + s < n) red[tid] += red[tid + s];
        __syncthreads();
    }
    double r = red[0];
    __syncthreads();
    return r;
}

// ------------------------------------------------------------------------------------------
// Block -> (kz tile, row) mapping for the x passes.  When a tile row is narrower than a 128-byte
// line (PAIR), the two tiles 2j, 2j+1 that share each line are given linear block ids L and L+8:
// blocks are dispatched round-robin over the 8 XCDs, so both halves of every line are fetched
// through the same XCD's L2 one dispatch round apart.  1-D grid of 16*ceil(npairs*nrows/8) blocks.
// ------------------------------------------------------------------------------------------
template <bool PAIR>
__device__ __forceinline__ bool tile_of_block(int ntiles, int nrows, int& tile, int& row) {
    if constexpr (!PAIR) {
        tile = blockIdx.x;
        row = blockIdx.y;
        return true;
    } else {
        const int L = blockIdx.x;
        const int grp = L >> 4, r = L & 15;
        const int q = grp * 8 + (r & 7);
        const int npairs = ntiles >> 1;
        if (q >= npairs * nrows) return false;
        tile = 2 * (q % npairs) + (r >> 3);
        row = q / npairs;
        return true;
    }
}

// ------------------------------------------------------------------------------------------
// Pass 1: stage input, curl (Solver/src/solver.c:637-650) and inverse x transform.
// grid (NZP/W, NK); block W*TPF; thread (t, w) = (tid / W, tid % W)
// fields: sel 0 -> (u, w) six fields; sel 1 -> u only; sel 2 -> w only (3 fields)
// ------------------------------------------------------------------------------------------
// batches: box = blockIdx.z; per-box element strides sbox (state/K) and obox (A1)
struct P1Args {
    const double2* uh;
    const void* kprev;  // RK4 schemes: K_{s-1}
    StageRow row;       // Dormand-Prince schemes: the stage's slopes and tableau row
    void* out;
    const void* twN;
    const DevScalars* sc;
    double* fslots;
    double kf2;
    long long sbox, obox;
    int visc_pow, scheme, stage, force, sel;
    int kmax, NK, NZP;
};

// FAM 0: RK4 schemes; FAM 1: Dormand-Prince family (separate code objects, so the RK4 kernels
// keep their register allocation)
template <typename T, int N, int EPT, int W, int FAM>
__global__ void __launch_bounds__(N / EPT * W) pass1_kernel(P1Args a) {
    using C = cpx_t<T>;
    constexpr int TPF = N / EPT;
    constexpr bool PAIR = W * sizeof(C) < 128;
    __shared__ C lds[N * W];
    __shared__ double red[N / EPT * W];
    int tile, iy;
    // compile-time extents: measured faster for the FP32 instance (0.69 -> 0.62 ms at 512^3) and
    // slower for FP64 (0.94 -> 1.00 ms), so only the FP32 instance takes them
    constexpr bool CX = !std::is_same<T, double>::value;
    const int kmax = CX ? Ext<N>::kmax : a.kmax, NK = CX ? Ext<N>::NK : a.NK, NZP = CX ? Ext<N>::NZP : a.NZP;
    if (!tile_of_block<PAIR>(NZP / W, NK, tile, iy)) return;
    const int box = blockIdx.z;
    const double2* __restrict__ uh = a.uh + box * a.sbox;
    const cpx_t<T>* kprev = a.kprev ? (const cpx_t<T>*)a.kprev + box * a.sbox : nullptr;
    const DevScalars& sc = a.sc[box];
    const int tid = threadIdx.x, w = tid % W, t = tid / W;
    const int kz = tile * W + w;
    const int ky = k_of_cidx(iy, kmax, NK);
    const bool colok = kz <= kmax;
    const double dt = sc.dt, nu = sc.nu;
    const long long plane = (long long)NK * NK * NZP;

    C u[3][EPT];  // stage input in the bulk precision (formed in FP64)
    double fpart = 0.0;
    if constexpr (FAM == 1) {
#pragma unroll
        for (int m = 0; m < EPT; ++m) {
            const int p = t + TPF * m;
            const int kx = k_of_pos(p, N);
            const bool ok = colok && kx >= -kmax && kx <= kmax;
            const long long idx = sidx(cidx_of_k(kx, NK), iy, kz, NK, NZP);
            const long long k2 = (long long)kx * kx + (long long)ky * ky + (long long)kz * kz;
            double2 ud[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) ud[c] = ok ? uh[c * plane + idx] : make_double2(0.0, 0.0);
            if (a.stage > 1 && ok) add_slopes<T>(ud, a.row, a.stage - 1, box * a.sbox, idx, plane, dt);
            if (a.force && ok && k2 > 0 && (double)k2 <= a.kf2) {
                const double wgt = (kz == 0) ? 1.0 : 2.0;
                fpart += wgt * (ud[0].x * ud[0].x + ud[0].y * ud[0].y + ud[1].x * ud[1].x + ud[1].y * ud[1].y +
                                ud[2].x * ud[2].x + ud[2].y * ud[2].y);
            }
#pragma unroll
            for (int c = 0; c < 3; ++c) u[c][m] = from_d2<T>(ud[c]);
        }
    } else {
        // RK4 family: every read of a round is issued before any is used (unconditional reads,
        // positions outside the cut read element 0 of the plane and are zeroed afterwards), so they
        // are in flight together instead of one position at a time; slots whose positions all lie
        // outside the cut (|kx| > N/3 for every thread) are neither read nor transformed as data.
        // FP32 takes two rounds of half the positions (register budget of a 1024-thread block).
        constexpr int KM = Ext<N>::kmax;
        constexpr int NH = (sizeof(T) == 4 && EPT % 2 == 0) ? 2 : 1, MH = EPT / NH;
        const bool prev = a.stage > 1;
        const bool rkcn = a.scheme == RK4CN;
        const double be_cn = (a.stage == 4) ? dt : 0.5 * dt;
#pragma unroll
        for (int h = 0; h < NH; ++h) {
            double2 ud[3][MH];
            C kv[3][MH];
#pragma unroll
            for (int j = 0; j < MH; ++j) {
                const int m = h * MH + j;
                if (TPF * m > KM && TPF * m + TPF - 1 < N - KM) continue;  // folded once unrolled
                const int kx = k_of_pos(t + TPF * m, N);
                const bool ok = colok && kx >= -kmax && kx <= kmax;
                const long long idx = ok ? sidx(cidx_of_k(kx, NK), iy, kz, NK, NZP) : 0;
#pragma unroll
                for (int c = 0; c < 3; ++c) ud[c][j] = uh[c * plane + idx];
                if (prev) {
#pragma unroll
                    for (int c = 0; c < 3; ++c) kv[c][j] = kprev[c * plane + idx];
                }
            }
#pragma unroll
            for (int j = 0; j < MH; ++j) {
                const int m = h * MH + j;
                if (TPF * m > KM && TPF * m + TPF - 1 < N - KM) {
#pragma unroll
                    for (int c = 0; c < 3; ++c) u[c][m] = cmk<C>(0, 0);
                    continue;
                }
                const int kx = k_of_pos(t + TPF * m, N);
                const bool ok = colok && kx >= -kmax && kx <= kmax;
                const long long k2 = (long long)kx * kx + (long long)ky * ky + (long long)kz * kz;
                double2 ud3[3];
#pragma unroll
                for (int c = 0; c < 3; ++c) ud3[c] = ok ? ud[c][j] : make_double2(0.0, 0.0);
                if (prev && ok) {
                    double al = 1.0, be = be_cn;
                    if (!rkcn) stage_in_coef(a.scheme, a.stage, visc_rate(nu, a.visc_pow, k2), dt, al, be);
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        const double2 kd = to_d2<T>(kv[c][j]);
                        ud3[c] = make_double2(al * ud3[c].x + be * kd.x, al * ud3[c].y + be * kd.y);
                    }
                }
                if (a.force && ok && k2 > 0 && (double)k2 <= a.kf2) {
                    const double wgt = (kz == 0) ? 1.0 : 2.0;
                    fpart += wgt * (ud3[0].x * ud3[0].x + ud3[0].y * ud3[0].y + ud3[1].x * ud3[1].x +
                                    ud3[1].y * ud3[1].y + ud3[2].x * ud3[2].x + ud3[2].y * ud3[2].y);
                }
#pragma unroll
                for (int c = 0; c < 3; ++c) u[c][m] = from_d2<T>(ud3[c]);
            }
        }
    }
    if (a.force && tile == 0 && (double)ky * ky <= a.kf2) {
        double s = block_sum(fpart, red);
        if (tid == 0) a.fslots[box * NK + iy] = s;
    }

    const int f0 = (a.sel == 2) ? 3 : 0;
    const int f1 = (a.sel == 1) ? 3 : 6;
    C* out = (C*)a.out + box * a.obox;
    const C* tw = (const C*)a.twN;
    const long long oplane = (long long)N * xplane(NK, NZP);
    // fully unrolled: with a runtime field index u[f][m] would live in scratch memory (measured:
    // 1.5x the A1 traffic in both reads and writes)
#pragma unroll
    for (int f = 0; f < 6; ++f) {
        if (f < f0 || f >= f1) continue;
        C v[EPT];
        using R = T;
#pragma unroll
        for (int m = 0; m < EPT; ++m) {
            if (f < 3) {
                v[m] = u[f][m];
            } else {
                const R kx = (R)k_of_pos(t + TPF * m, N), kyd = (R)ky, kzd = (R)kz;
                // w = i k x u
                R xr, xi;
                if (f == 3) { xr = kyd * u[2][m].x - kzd * u[1][m].x; xi = kyd * u[2][m].y - kzd * u[1][m].y; }
                else if (f == 4) { xr = kzd * u[0][m].x - kx * u[2][m].x; xi = kzd * u[0][m].y - kx * u[2][m].y; }
                else { xr = kx * u[1][m].x - kyd * u[0][m].x; xi = kx * u[1][m].y - kyd * u[0][m].y; }
                v[m] = cmk<C>(-xi, xr);
            }
        }
        fft<N, EPT, +1>(v, t, lds, w, W, tw);
        const int fo = f - f0;
#pragma unroll
        for (int m = 0; m < EPT; ++m) {
            const int p = t + TPF * m;
            out[fo * oplane + xidx<N>(p, iy, kz, NK, NZP)] = v[m];
        }
    }
}

// ------------------------------------------------------------------------------------------
// Passes 2 and 4: column transforms along y.
//   DIR=+1 (pass 2): rows in = compact ky (NK), rows out = all N positions
//   DIR=-1 (pass 4): rows in = all N positions, rows out = compact ky
// in/out layouts [F][N(x)][rows][NZP]; grid (NZP/W, N, F)
// ------------------------------------------------------------------------------------------
// batches: blockIdx.z = box * F + field; per-box element strides ibox, obox
struct PYArgs {
    const void* in;
    void* out;
    const void* twN;
    long long ibox, obox;
    int F;
    int kmax, NK, NZP;
};

template <typename T, int N, int EPT, int W, int DIR>
__global__ void __launch_bounds__(N / EPT * W) passy_kernel(PYArgs a) {
    using C = cpx_t<T>;
    constexpr int TPF = N / EPT;
    __shared__ C lds[N * W];
    const int tid = threadIdx.x, w = tid % W, t = tid / W;
    const int kz = blockIdx.x * W + w;
    const int x = blockIdx.y, f = blockIdx.z % a.F, box = blockIdx.z / a.F;
    const int kmax = a.kmax, NK = a.NK, NZP = a.NZP;
    // DIR>0: in = A1, out = A2;  DIR<0: in = A3, out = A4 (A1/A4 planes padded by XPAD)
    const long long fin = (DIR > 0) ? (long long)N * xplane(NK, NZP) : (long long)N * N * NZP;
    const long long fout = (DIR > 0) ? (long long)N * N * NZP : (long long)N * xplane(NK, NZP);
    const C* in = (const C*)a.in + box * a.ibox + (long long)f * fin;
    C* out = (C*)a.out + box * a.obox + (long long)f * fout;
    C v[EPT];
#pragma unroll
    for (int m = 0; m < EPT; ++m) {
        const int p = t + TPF * m;
        if constexpr (DIR > 0) {
            const int k = k_of_pos(p, N);
            const bool ok = k >= -kmax && k <= kmax;
            v[m] = ok ? in[xidx<N>(x, cidx_of_k(k, NK), kz, NK, NZP)] : cmk<C>(0, 0);
        } else {
            v[m] = in[((long long)x * N + p) * NZP + kz];
        }
    }
    fft<N, EPT, DIR>(v, t, lds, w, W, (const C*)a.twN);
#pragma unroll
    for (int m = 0; m < EPT; ++m) {
        const int p = t + TPF * m;
        if constexpr (DIR > 0) {
            out[((long long)x * N + p) * NZP + kz] = v[m];
        } else {
            const int k = k_of_pos(p, N);
            if (k >= -kmax && k <= kmax) out[xidx<N>(x, cidx_of_k(k, NK), kz, NK, NZP)] = v[m];
        }
    }
}

// ------------------------------------------------------------------------------------------
// Pass 3: per (x,y) line along z.
//  mode ZM_RHS : c2r of the six fields (u, w), u x w in registers (Solver/src/solver.c:664-677),
//                r2c of the three products; also the running max of |ux|+|uy|+|uz|
//  mode ZM_REAL: c2r of three fields, stored as real AoS [x][y][z][3] scaled by `scale`
//  mode ZM_R2C : r2c of a real AoS [x][y][z][3] field
// Half-length transforms (M = N/2): c2r forms Z[k] = (X[k] + X*[M-k]) + i e^{2pi i k/N}(X[k] - X*[M-k])
// and z = IDFT_M(Z) gives x[2n] = Re z[n], x[2n+1] = Im z[n]; r2c packs z[n] = x[2n] + i x[2n+1],
// Z = DFT_M(z), X[k] = (Z[k] + Z*[M-k])/2 - (i/2) e^{-2pi i k/N} (Z[k] - Z*[M-k]).
// grid ceil(N*N/LPB); block LPB*TPFZ
// ------------------------------------------------------------------------------------------
enum ZMode { ZM_RHS = 0, ZM_REAL = 1, ZM_R2C = 2 };

// batches: lines of box b are lids [b N^2, (b+1) N^2); per-box element strides ibox, obox
struct PZArgs {
    const void* in;   // A2 (RHS, REAL) or real AoS doubles (R2C)
    void* out;        // A3 (RHS, R2C) or real AoS doubles (REAL)
    const void* twM;  // length N/2 table
    const void* twN;  // length N table
    DevScalars* sc;
    double scale;
    long long ibox, obox;
    int nbox;
    int kmax, NZP;
    int track_max;
    int soa;  // ZM_REAL only: 1 writes the components as three planes [3][x][y][z]
};

template <int M, int EPTZ> constexpr int zsync() { return (M / EPTZ <= 64 && 64 % (M / EPTZ) == 0) ? SYNC_WAVE : SYNC_BLOCK; }
template <int M, int EPTZ> constexpr int zlds() { return fft_lds_len(M, EPTZ, true); }

template <typename T, int N, int EPTZ>
__device__ __forceinline__ void z_c2r_field(cpx_t<T> (&v)[EPTZ], const cpx_t<T>* line, int kmax, int t,
                                            cpx_t<T>* lds, int base, const cpx_t<T>* twM, const cpx_t<T>* twN) {
    using C = cpx_t<T>;
    constexpr int M = N / 2, TPF = M / EPTZ;
#pragma unroll
    for (int m = 0; m < EPTZ; ++m) {
        const int k = t + TPF * m;
        const int kr = M - k;
        C xk = (k <= kmax) ? line[k] : cmk<C>(0, 0);
        C xr = (kr <= kmax) ? line[kr] : cmk<C>(0, 0);
        if (k == 0) { xk.y = 0; xr.y = 0; }  // c2r uses only the real parts of X[0], X[N/2]
        const C xrc = cconj(xr);
        const C s = cadd(xk, xrc), d = csub(xk, xrc);
        C e = twN[k];
        e.y = -e.y;  // e^{+2 pi i k / N}
        v[m] = cadd(s, cmul_i<+1>(cmul(e, d)));
    }
    fft<M, EPTZ, +1, zsync<M, EPTZ>(), true>(v, t, lds, base, 1, twM);
}

template <typename T, int N, int EPTZ>
__device__ __forceinline__ void z_r2c_store(cpx_t<T> (&v)[EPTZ], cpx_t<T>* outline, int kmax, int NZP, int t,
                                            cpx_t<T>* lds, int base, const cpx_t<T>* twM, const cpx_t<T>* twN,
                                            bool live) {
    using C = cpx_t<T>;
    using R = decltype(C::x);
    constexpr int M = N / 2, TPF = M / EPTZ;
    constexpr int SY = zsync<M, EPTZ>();
    fft<M, EPTZ, -1, SY, true>(v, t, lds, base, 1, twM);
#pragma unroll
    for (int m = 0; m < EPTZ; ++m) lds[base + t + TPF * m] = v[m];
    lds_sync<SY>();
#pragma unroll
    for (int m = 0; m < EPTZ; ++m) {
        const int k = t + TPF * m;
        if (live && k < NZP) {
            C r = cmk<C>(0, 0);
            if (k <= kmax) {
                const C zk = v[m];
                const C zr = cconj(lds[base + ((M - k) % M)]);
                const C s = cadd(zk, zr), d = csub(zk, zr);
                const C e = twN[k];  // e^{-2 pi i k / N}
                const C ed = cmul(e, d);
                // X = s/2 - (i/2) e d
                r = cmk<C>(R(0.5) * (s.x + ed.y), R(0.5) * (s.y - ed.x));
            }
            outline[k] = r;
        }
    }
    if (live)
        for (int k = M + t; k < NZP; k += TPF) outline[k] = cmk<C>(0, 0);
    lds_sync<SY>();
}

template <typename T, int N, int EPTZ, int LPB, int MODE>
__global__ void __launch_bounds__(N / 2 / EPTZ * LPB) passz_kernel(PZArgs a) {
    using C = cpx_t<T>;
    using R = decltype(C::x);
    constexpr int M = N / 2, TPF = M / EPTZ;
    constexpr int MP = zlds<M, EPTZ>();
    __shared__ C lds[MP * LPB];
    const int tid = threadIdx.x, line = tid / TPF, t = tid % TPF;
    long long gid = (long long)blockIdx.x * LPB + line;
    const long long nl = (long long)N * N;
    const bool live = gid < nl * a.nbox;
    if (!live) gid = nl * a.nbox - 1;  // idle lines recompute the last line but never store
    const int box = (int)(gid / nl);
    const long long lid = gid - box * nl;
    const int base = line * MP;
    const int kmax = a.kmax, NZP = a.NZP;
    const C* twM = (const C*)a.twM;
    const C* twN = (const C*)a.twN;
    const long long fstride = (long long)N * N * NZP;

    if constexpr (MODE == ZM_RHS) {
        const C* in = (const C*)a.in + box * a.ibox + lid * NZP;
        C* out = (C*)a.out + box * a.obox + lid * NZP;
        C uu[3][EPTZ], ww[3][EPTZ];
#pragma unroll
        for (int c = 0; c < 3; ++c) z_c2r_field<T, N, EPTZ>(uu[c], in + c * fstride, kmax, t, lds, base, twM, twN);
#pragma unroll
        for (int c = 0; c < 3; ++c) z_c2r_field<T, N, EPTZ>(ww[c], in + (3 + c) * fstride, kmax, t, lds, base, twM, twN);
        if (a.track_max) {
            double mx = 0.0;
#pragma unroll
            for (int m = 0; m < EPTZ; ++m) {
                double s0 = fabs((double)uu[0][m].x) + fabs((double)uu[1][m].x) + fabs((double)uu[2][m].x);
                double s1 = fabs((double)uu[0][m].y) + fabs((double)uu[1][m].y) + fabs((double)uu[2][m].y);
                mx = nanmax(mx, nanmax(s0, s1));
            }
            mx *= a.scale;
            // a wavefront's lines are consecutive and N^2 >= 64, so they span at most two boxes
            // (b0 and b0 + 1) when N^2 is not a multiple of the lines per wavefront
            const int b0 = __shfl(box, 0);
            double ma = (box == b0) ? mx : 0.0, mb = (box == b0) ? 0.0 : mx;
            for (int off = 32; off > 0; off >>= 1) {
                ma = nanmax(ma, __shfl_xor(ma, off));
                mb = nanmax(mb, __shfl_xor(mb, off));
            }
            if ((tid & 63) == 0) {
                atomicMax(&a.sc[b0].maxbits, (unsigned long long)__double_as_longlong(ma));
                if (b0 + 1 < a.nbox) atomicMax(&a.sc[b0 + 1].maxbits, (unsigned long long)__double_as_longlong(mb));
            }
        }
        C pr[3][EPTZ];
#pragma unroll
        for (int m = 0; m < EPTZ; ++m) {
            // real and imaginary parts are two independent grid points
            pr[0][m] = cmk<C>(uu[1][m].x * ww[2][m].x - uu[2][m].x * ww[1][m].x, uu[1][m].y * ww[2][m].y - uu[2][m].y * ww[1][m].y);
            pr[1][m] = cmk<C>(uu[2][m].x * ww[0][m].x - uu[0][m].x * ww[2][m].x, uu[2][m].y * ww[0][m].y - uu[0][m].y * ww[2][m].y);
            pr[2][m] = cmk<C>(uu[0][m].x * ww[1][m].x - uu[1][m].x * ww[0][m].x, uu[0][m].y * ww[1][m].y - uu[1][m].y * ww[0][m].y);
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) z_r2c_store<T, N, EPTZ>(pr[c], out + c * fstride, kmax, NZP, t, lds, base, twM, twN, live);
    } else if constexpr (MODE == ZM_REAL) {
        const C* in = (const C*)a.in + box * a.ibox + lid * NZP;
        double* out = (double*)a.out + box * a.obox + lid * (long long)N * 3;
        const long long plane3 = (long long)N * N * N;
        double* outs = (double*)a.out + box * a.obox + lid * (long long)N;  // SoA: line lid of plane c
        for (int c = 0; c < 3; ++c) {
            C v[EPTZ];
            z_c2r_field<T, N, EPTZ>(v, in + c * fstride, kmax, t, lds, base, twM, twN);
            if (live) {
#pragma unroll
                for (int m = 0; m < EPTZ; ++m) {
                    const int n = t + TPF * m;
                    if (a.soa) {
                        ((double2*)(outs + c * plane3))[n] = make_double2((double)v[m].x * a.scale, (double)v[m].y * a.scale);
                    } else {
                        out[(2 * n) * 3 + c] = (double)v[m].x * a.scale;
                        out[(2 * n + 1) * 3 + c] = (double)v[m].y * a.scale;
                    }
                }
            }
        }
    } else {  // ZM_R2C
        const double* in = (const double*)a.in + box * a.ibox + lid * (long long)N * 3;
        C* out = (C*)a.out + box * a.obox + lid * NZP;
        for (int c = 0; c < 3; ++c) {
            C v[EPTZ];
#pragma unroll
            for (int m = 0; m < EPTZ; ++m) {
                const int n = t + TPF * m;
                v[m] = cmk<C>((R)in[(2 * n) * 3 + c], (R)in[(2 * n + 1) * 3 + c]);
            }
            z_r2c_store<T, N, EPTZ>(v, out + c * fstride, kmax, NZP, t, lds, base, twM, twN, live);
        }
    }
}

// ------------------------------------------------------------------------------------------
// Pass 3, RHS form, persistent per-wavefront version.  Each wavefront owns 64/TPF lines at a
// time and loops over line groups; all LDS exchanges are wavefront-local.  Per line: six
// half-length c2r transforms (u then w), the product u x w accumulated field by field
// (P_x = u_y w_z - u_z w_y, ...), then three r2c transforms.  The coefficients X[k] of field f+1
// are loaded while field f is transformed; the partner X[M-k] comes from LDS.  Twiddle tables
// live in LDS (one copy per block).
// ------------------------------------------------------------------------------------------
template <typename T, int N, int EPTZ>
__device__ __forceinline__ void zr_load(cpx_t<T> (&x)[EPTZ], const cpx_t<T>* line, int kmax, int t) {
    constexpr int TPF = (N / 2) / EPTZ;
#pragma unroll
    for (int m = 0; m < EPTZ; ++m) {
        const int k = t + TPF * m;
        x[m] = (k <= kmax) ? line[k] : cmk<cpx_t<T>>(0, 0);
    }
}

// c2r of one field whose X[k] (k = t + TPF m) are in x; result z[n] = x[2n] + i x[2n+1] in v
template <int M, int EPTZ> constexpr int zntw() { return num_base_twiddles(M, EPTZ) > 0 ? num_base_twiddles(M, EPTZ) : 1; }
// twiddle registers of the half-length transform: base twiddles (powers formed by multiplication)
// or, with TAB, the full table of every butterfly's twiddles
template <int M, int EPTZ, bool TAB> constexpr int zntw2() { return TAB ? ntab(M, EPTZ) : zntw<M, EPTZ>(); }

template <int M, int EPTZ, int DIR, int SY, bool TAB, typename C>
__device__ __forceinline__ void zfft(C (&v)[EPTZ], int t, C* lds, int base, const C (&w)[zntw2<M, EPTZ, TAB>()]) {
    if constexpr (TAB) fft_tab<M, EPTZ, DIR, SY, true>(v, t, lds, base, 1, w);
    else fft_reg<M, EPTZ, DIR, SY, true, zntw<M, EPTZ>()>(v, t, lds, base, 1, w);
}

template <typename T, int N, int EPTZ, int SY, bool TAB = false>
__device__ __forceinline__ void zr_c2r(cpx_t<T> (&v)[EPTZ], const cpx_t<T> (&x)[EPTZ], int t, cpx_t<T>* lds, int base,
                                       const cpx_t<T> (&wz)[zntw2<N / 2, EPTZ, TAB>()], const cpx_t<T> (&ek)[EPTZ]) {
    using C = cpx_t<T>;
    constexpr int M = N / 2, TPF = M / EPTZ;
#pragma unroll
    for (int m = 0; m < EPTZ; ++m) lds[base + t + TPF * m] = x[m];
    lds_sync<SY>();
#pragma unroll
    for (int m = 0; m < EPTZ; ++m) {
        const int k = t + TPF * m;
        C xk = x[m];
        C xr = lds[base + ((M - k) & (M - 1))];  // X[M-k]; for k = 0 this is X[0], replaced below
        if (k == 0) { xk.y = 0; xr = cmk<C>(0, 0); }  // X[N/2] is outside the 2/3 box: zero
        // Z = s + i e^{+2 pi i k/N} d with s = X[k] + X*[M-k], d = X[k] - X*[M-k]; written out so
        // that each component is two FMAs on s (i conj(ek) = (ek.y, ek.x))
        using R = decltype(C::x);
        const R sx = xk.x + xr.x, sy = xk.y - xr.y, dx = xk.x - xr.x, dy = xk.y + xr.y;
        const C e = ek[m];  // e^{-2 pi i k / N}
        v[m] = cmk<C>(sx + e.y * dx - e.x * dy, sy + e.y * dy + e.x * dx);
    }
    lds_sync<SY>();
    zfft<M, EPTZ, +1, SY, TAB>(v, t, lds, base, wz);
}

template <typename T, int N, int EPTZ, int SY, bool TAB = false>
__device__ __forceinline__ void zr_r2c_store(cpx_t<T> (&v)[EPTZ], cpx_t<T>* outline, int kmax, int NZP, int t,
                                             cpx_t<T>* lds, int base, const cpx_t<T> (&wz)[zntw2<N / 2, EPTZ, TAB>()],
                                             const cpx_t<T> (&ek)[EPTZ], bool live) {
    using C = cpx_t<T>;
    using R = decltype(C::x);
    constexpr int M = N / 2, TPF = M / EPTZ;
    zfft<M, EPTZ, -1, SY, TAB>(v, t, lds, base, wz);
#pragma unroll
    for (int m = 0; m < EPTZ; ++m) lds[base + t + TPF * m] = v[m];
    lds_sync<SY>();
#pragma unroll
    for (int m = 0; m < EPTZ; ++m) {
        const int k = t + TPF * m;
        if (live && k < NZP) {
            C r = cmk<C>(0, 0);
            if (k <= kmax) {
                // X = (s - i e^{-2 pi i k/N} d) / 2 with s = Z[k] + Z*[M-k], d = Z[k] - Z*[M-k]
                const C zk = v[m];
                const C zm = lds[base + ((M - k) & (M - 1))];
                const R sx = zk.x + zm.x, sy = zk.y - zm.y, dx = zk.x - zm.x, dy = zk.y + zm.y;
                const C e = ek[m];
                r = cmk<C>(R(0.5) * (sx + e.x * dy + e.y * dx), R(0.5) * (sy - e.x * dx + e.y * dy));
            }
            outline[k] = r;
        }
    }
    if (live)
        for (int k = M + t; k < NZP; k += TPF) outline[k] = cmk<C>(0, 0);
    lds_sync<SY>();
}

template <typename T, int N, int EPTZ, int WPB, int OCC>
__global__ void __launch_bounds__(64 * WPB, OCC) passz_rhs_kernel(PZArgs a) {
    using C = cpx_t<T>;
    constexpr int M = N / 2, TPF = M / EPTZ, LPW = 64 / TPF;
    static_assert(TPF <= 64 && 64 % TPF == 0 && (M & (M - 1)) == 0, "line must fit a wavefront, M power of two");
    constexpr int SY = SYNC_WAVE;
    constexpr int MP = zlds<M, EPTZ>();
    __shared__ C lds[WPB * LPW * MP];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int sub = lane / TPF, t = lane % TPF;
    // per-thread twiddles, fixed for every line this thread handles
    C wz[zntw<M, EPTZ>()];
    fft_load_base_twiddles<M, EPTZ>(wz, t, (const C*)a.twM);
    C ek[EPTZ];  // e^{-2 pi i k / N}, k = t + TPF m
#pragma unroll
    for (int m = 0; m < EPTZ; ++m) ek[m] = ((const C*)a.twN)[t + TPF * m];
    const int base = (wv * LPW + sub) * MP;
    constexpr int kmax = Ext<N>::kmax, NZP = Ext<N>::NZP;
    const long long nlines = (long long)N * N;        // lines per box (a multiple of LPW)
    const long long total = nlines * a.nbox;
    const long long ngroups = (total + LPW - 1) / LPW;
    const long long fstride = nlines * NZP;
    const C* in0 = (const C*)a.in;
    C* out0 = (C*)a.out;
    // line address of global line id gl (box-major)
    auto line_off = [&](long long gl, long long stride) {
        const long long b = gl / nlines;
        return b * stride + (gl - b * nlines) * NZP;
    };
    double mx = 0.0;
    int mx_box = -1;  // box whose maximum mx holds (uniform across the wavefront)
    auto flush_max = [&]() {
        if (a.track_max && mx_box >= 0) {
            double v = mx * a.scale;
            for (int off = 32; off > 0; off >>= 1) v = nanmax(v, __shfl_xor(v, off));
            if (lane == 0) atomicMax(&a.sc[mx_box].maxbits, (unsigned long long)__double_as_longlong(v));
        }
        mx = 0.0;
    };

    long long g = (long long)blockIdx.x * WPB + wv;
    C cur[EPTZ];
    if (g < ngroups) {
        long long gl = g * LPW + sub;
        if (gl >= total) gl = total - 1;
        zr_load<T, N, EPTZ>(cur, in0 + line_off(gl, a.ibox), kmax, t);
    }
    for (; g < ngroups; g += (long long)gridDim.x * WPB) {
        long long gl = g * LPW + sub;
        const bool live = gl < total;
        if (!live) gl = total - 1;  // idle lanes recompute the last line but never store
        const int box = (int)((g * LPW) / nlines);
        if (box != mx_box) {
            flush_max();
            mx_box = box;
        }
        const C* in = in0 + line_off(gl, a.ibox);
        C ur[3][EPTZ], pr[3][EPTZ];
#pragma unroll
        for (int f = 0; f < 6; ++f) {
            C nxt[EPTZ];
            if (f < 5) zr_load<T, N, EPTZ>(nxt, in + (f + 1) * fstride, kmax, t);
            C v[EPTZ];
            zr_c2r<T, N, EPTZ, SY>(v, cur, t, lds, base, wz, ek);
#pragma unroll
            for (int m = 0; m < EPTZ; ++m) {
                const C w = v[m];
                if (f < 3) {
                    ur[f][m] = w;
                } else if (f == 3) {  // w_x: P_y = u_z w_x, P_z = -u_y w_x
                    pr[1][m] = cmk<C>(ur[2][m].x * w.x, ur[2][m].y * w.y);
                    pr[2][m] = cmk<C>(-(ur[1][m].x * w.x), -(ur[1][m].y * w.y));
                } else if (f == 4) {  // w_y: P_x = -u_z w_y, P_z += u_x w_y
                    pr[0][m] = cmk<C>(-(ur[2][m].x * w.x), -(ur[2][m].y * w.y));
                    pr[2][m] = cmk<C>(pr[2][m].x + ur[0][m].x * w.x, pr[2][m].y + ur[0][m].y * w.y);
                } else {  // w_z: P_x += u_y w_z, P_y -= u_x w_z
                    pr[0][m] = cmk<C>(pr[0][m].x + ur[1][m].x * w.x, pr[0][m].y + ur[1][m].y * w.y);
                    pr[1][m] = cmk<C>(pr[1][m].x - ur[0][m].x * w.x, pr[1][m].y - ur[0][m].y * w.y);
                }
            }
            if (f < 5) {
#pragma unroll
                for (int m = 0; m < EPTZ; ++m) cur[m] = nxt[m];
            }
        }
        // prefetch field 0 of the next line group while the products are transformed back
        const long long gn = g + (long long)gridDim.x * WPB;
        if (gn < ngroups) {
            long long ln = gn * LPW + sub;
            if (ln >= total) ln = total - 1;
            zr_load<T, N, EPTZ>(cur, in0 + line_off(ln, a.ibox), kmax, t);
        }
        if (a.track_max && live) {
#pragma unroll
            for (int m = 0; m < EPTZ; ++m) {
                const double s0 = fabs((double)ur[0][m].x) + fabs((double)ur[1][m].x) + fabs((double)ur[2][m].x);
                const double s1 = fabs((double)ur[0][m].y) + fabs((double)ur[1][m].y) + fabs((double)ur[2][m].y);
                mx = nanmax(mx, nanmax(s0, s1));
            }
        }
        C* out = out0 + line_off(gl, a.obox);
#pragma unroll
        for (int c = 0; c < 3; ++c) zr_r2c_store<T, N, EPTZ, SY>(pr[c], out + c * fstride, kmax, NZP, t, lds, base, wz, ek, live);
    }
    flush_max();
}

// ------------------------------------------------------------------------------------------
// Pass 3, RHS form, reduced-liveness persistent version.  Same lines, exchanges and outputs as
// passz_rhs_kernel, but the six fields are transformed in the order u_y, w_z, u_z, w_y, u_x, w_x
// so that P_x = u_y w_z - u_z w_y is complete after four fields and is transformed back and stored
// before u_x arrives; then P_y = u_z w_x - u_x w_z and P_z = u_x w_y - u_y w_x.  At most six
// fields are live (three u/w fields, P_x, the transform in progress and the prefetched next field)
// instead of nine, which leaves room for the full twiddle table (TAB) and for more waves per SIMD.
// ------------------------------------------------------------------------------------------
template <typename T, int N, int EPTZ, int WPB, int OCC, bool TAB>
__global__ void __launch_bounds__(64 * WPB, OCC) passz_rhs2_kernel(PZArgs a) {
    using C = cpx_t<T>;
    constexpr int M = N / 2, TPF = M / EPTZ, LPW = 64 / TPF;
    static_assert(TPF <= 64 && 64 % TPF == 0 && (M & (M - 1)) == 0, "line must fit a wavefront, M power of two");
    constexpr int SY = SYNC_WAVE;
    constexpr int MP = zlds<M, EPTZ>();
    __shared__ C lds[WPB * LPW * MP];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int sub = lane / TPF, t = lane % TPF;
    C wz[zntw2<M, EPTZ, TAB>()];
    if constexpr (TAB) fft_load_twiddle_table<M, EPTZ>(wz, t, (const C*)a.twM);
    else fft_load_base_twiddles<M, EPTZ>(wz, t, (const C*)a.twM);
    C ek[EPTZ];  // e^{-2 pi i k / N}, k = t + TPF m
#pragma unroll
    for (int m = 0; m < EPTZ; ++m) ek[m] = ((const C*)a.twN)[t + TPF * m];
    const int base = (wv * LPW + sub) * MP;
    constexpr int kmax = Ext<N>::kmax, NZP = Ext<N>::NZP;
    const long long nlines = (long long)N * N;  // lines per box (a multiple of LPW)
    const long long total = nlines * a.nbox;
    const long long ngroups = (total + LPW - 1) / LPW;
    const long long fstride = nlines * NZP;
    const C* in0 = (const C*)a.in;
    C* out0 = (C*)a.out;
    auto line_off = [&](long long gl, long long stride) {
        const long long b = gl / nlines;
        return b * stride + (gl - b * nlines) * NZP;
    };
    double mx = 0.0;
    int mx_box = -1;
    auto flush_max = [&]() {
        if (a.track_max && mx_box >= 0) {
            double v = mx * a.scale;
            for (int off = 32; off > 0; off >>= 1) v = nanmax(v, __shfl_xor(v, off));
            if (lane == 0) atomicMax(&a.sc[mx_box].maxbits, (unsigned long long)__double_as_longlong(v));
        }
        mx = 0.0;
    };

    long long g = (long long)blockIdx.x * WPB + wv;
    C cur[EPTZ];  // coefficients of the field transformed next
    if (g < ngroups) {
        long long gl = g * LPW + sub;
        if (gl >= total) gl = total - 1;
        zr_load<T, N, EPTZ>(cur, in0 + line_off(gl, a.ibox) + 1 * fstride, kmax, t);  // u_y
    }
    for (; g < ngroups; g += (long long)gridDim.x * WPB) {
        long long gl = g * LPW + sub;
        const bool live = gl < total;
        if (!live) gl = total - 1;  // idle lanes recompute the last line but never store
        const int box = (int)((g * LPW) / nlines);
        if (box != mx_box) {
            flush_max();
            mx_box = box;
        }
        const C* in = in0 + line_off(gl, a.ibox);
        C* out = out0 + line_off(gl, a.obox);
        C nxt[EPTZ], v[EPTZ], uy[EPTZ], wzf[EPTZ], uz[EPTZ], wy[EPTZ], px[EPTZ], py[EPTZ], pz[EPTZ];
        // u_y
        zr_load<T, N, EPTZ>(nxt, in + 5 * fstride, kmax, t);
        zr_c2r<T, N, EPTZ, SY, TAB>(uy, cur, t, lds, base, wz, ek);
        // w_z: P_x = u_y w_z
        zr_load<T, N, EPTZ>(cur, in + 2 * fstride, kmax, t);
        zr_c2r<T, N, EPTZ, SY, TAB>(wzf, nxt, t, lds, base, wz, ek);
#pragma unroll
        for (int m = 0; m < EPTZ; ++m) px[m] = cmk<C>(uy[m].x * wzf[m].x, uy[m].y * wzf[m].y);
        // u_z
        zr_load<T, N, EPTZ>(nxt, in + 4 * fstride, kmax, t);
        zr_c2r<T, N, EPTZ, SY, TAB>(uz, cur, t, lds, base, wz, ek);
        // w_y: P_x -= u_z w_y, complete
        zr_load<T, N, EPTZ>(cur, in + 0 * fstride, kmax, t);
        zr_c2r<T, N, EPTZ, SY, TAB>(wy, nxt, t, lds, base, wz, ek);
#pragma unroll
        for (int m = 0; m < EPTZ; ++m) px[m] = cmk<C>(px[m].x - uz[m].x * wy[m].x, px[m].y - uz[m].y * wy[m].y);
        zr_r2c_store<T, N, EPTZ, SY, TAB>(px, out + 0 * fstride, kmax, NZP, t, lds, base, wz, ek, live);
        // u_x: P_y = -u_x w_z, P_z = u_x w_y; the running max of |u_x| + |u_y| + |u_z|
        zr_load<T, N, EPTZ>(nxt, in + 3 * fstride, kmax, t);
        zr_c2r<T, N, EPTZ, SY, TAB>(v, cur, t, lds, base, wz, ek);
#pragma unroll
        for (int m = 0; m < EPTZ; ++m) {
            py[m] = cmk<C>(-(v[m].x * wzf[m].x), -(v[m].y * wzf[m].y));
            pz[m] = cmk<C>(v[m].x * wy[m].x, v[m].y * wy[m].y);
        }
        if (a.track_max && live) {
#pragma unroll
            for (int m = 0; m < EPTZ; ++m) {
                const double s0 = fabs((double)v[m].x) + fabs((double)uy[m].x) + fabs((double)uz[m].x);
                const double s1 = fabs((double)v[m].y) + fabs((double)uy[m].y) + fabs((double)uz[m].y);
                mx = nanmax(mx, nanmax(s0, s1));
            }
        }
        // w_x: P_y += u_z w_x, P_z -= u_y w_x; meanwhile u_y of the next line group is fetched
        const long long gn = g + (long long)gridDim.x * WPB;
        if (gn < ngroups) {
            long long ln = gn * LPW + sub;
            if (ln >= total) ln = total - 1;
            zr_load<T, N, EPTZ>(cur, in0 + line_off(ln, a.ibox) + 1 * fstride, kmax, t);
        }
        zr_c2r<T, N, EPTZ, SY, TAB>(v, nxt, t, lds, base, wz, ek);
#pragma unroll
        for (int m = 0; m < EPTZ; ++m) {
            py[m] = cmk<C>(py[m].x + uz[m].x * v[m].x, py[m].y + uz[m].y * v[m].y);
            pz[m] = cmk<C>(pz[m].x - uy[m].x * v[m].x, pz[m].y - uy[m].y * v[m].y);
        }
        zr_r2c_store<T, N, EPTZ, SY, TAB>(py, out + 1 * fstride, kmax, NZP, t, lds, base, wz, ek, live);
        zr_r2c_store<T, N, EPTZ, SY, TAB>(pz, out + 2 * fstride, kmax, NZP, t, lds, base, wz, ek, live);
    }
    flush_max();
}

// ------------------------------------------------------------------------------------------
// Pass 3, RHS form, full-length paired version.  Two real fields share one complex transform of
// length N: the inverse of Z = U_h + i W_h (U_h, W_h the Hermitian extensions of two half spectra,
// U_h[N - k] = conj U[k]) is u + i w, and the forward transform Z of p + i q gives
// P[k] = (Z[k] + conj Z[N - k]) / 2, Q[k] = (Z[k] - conj Z[N - k]) / (2 i).  Per line: three
// inverse transforms (u_y, w_z), (u_z, w_y), (u_x, w_x) -- each pair's product terms are formed as
// soon as it arrives -- and two forward transforms, (P_x, P_y) and P_z alone.  Against the
// half-length form this removes the pre/post rotation of every field and its LDS exchange: 5
// transforms of N points and 10 exchanges per line instead of 9 of N/2 points and 27 (radix 8 at
// N = 512).  Only kz <= kmax is nonzero on input, so positions kmax < p < N - kmax are zero; the
// element slots that are zero for every thread are compile-time zeros, and outputs beyond kmax
// are never formed.  The mirrored coefficients X[N - p] are read from global memory (the line was
// just fetched, so they come from cache); the separation partner Z[N - k] comes from the thread
// holding position N - k by a cross-lane shuffle (thread TPF - t, element EPT - 1 - m).
// One line per TPF = N / EPT threads (a power of two: TPF - t is the partner thread); lines per
// wavefront LPW = 64 / TPF; persistent loop.  N may carry a factor 3 (radix-3 stages).
// ------------------------------------------------------------------------------------------
// raw coefficients of one field for the positions p = t + TPF m of a full-length transform:
// X[p] for p <= kmax, X[N - p] for p >= N - kmax, zero between (compile-time zero where the whole
// element slot lies between)
template <typename T, int N, int EPT>
__device__ __forceinline__ void zf_load(cpx_t<T> (&x)[EPT], const cpx_t<T>* line, int t) {
    using C = cpx_t<T>;
    constexpr int TPF = N / EPT, kmax = Ext<N>::kmax;
#pragma unroll
    for (int m = 0; m < EPT; ++m) {
        if (TPF * m > kmax && TPF * m + TPF - 1 < N - kmax) {  // folded once the loop is unrolled
            x[m] = cmk<C>(0, 0);
        } else {
            const int p = t + TPF * m;
            const int src = (p <= kmax) ? p : ((p >= N - kmax) ? N - p : -1);
            x[m] = (src >= 0) ? line[src] : cmk<C>(0, 0);
        }
    }
}
// Z[p] = U_h[p] + i W_h[p] from the raw coefficients (conjugated on the mirrored side; the
// imaginary part of X[0] is dropped, as c2r does)
template <typename T, int N, int EPT>
__device__ __forceinline__ void zf_pack(cpx_t<T> (&z)[EPT], const cpx_t<T> (&u)[EPT], const cpx_t<T> (&w)[EPT], int t) {
    using C = cpx_t<T>;
    constexpr int TPF = N / EPT, kmax = Ext<N>::kmax;
#pragma unroll
    for (int m = 0; m < EPT; ++m) {
        if (TPF * m > kmax && TPF * m + TPF - 1 < N - kmax) {  // folded once the loop is unrolled
            z[m] = cmk<C>(0, 0);
        } else {
            const int p = t + TPF * m;
            C a = u[m], b = w[m];
            if (p > N / 2) { a.y = -a.y; b.y = -b.y; }
            if (p == 0) { a.y = 0; b.y = 0; }
            z[m] = cmk<C>(a.x - b.y, a.y + b.x);
        }
    }
}
template <int N, int EPT> constexpr int zf_mout() { return Ext<N>::kmax / (N / EPT) + 1; }  // slots reaching k <= kmax

// forward transform of p + i q (pair) or of p alone (!PAIR), then the stores of kz < NZP
template <typename T, int N, int EPT, int SY, bool TAB, bool PAIR>
__device__ __forceinline__ void zf_forward_store(cpx_t<T> (&z)[EPT], cpx_t<T>* out1, cpx_t<T>* out2, int t, int lane,
                                                 cpx_t<T>* lds, int base, const cpx_t<T> (&wz)[zntw2<N, EPT, TAB>()],
                                                 bool live, bool live2) {
    using C = cpx_t<T>;
    using R = decltype(C::x);
    constexpr int TPF = N / EPT, kmax = Ext<N>::kmax, NZP = Ext<N>::NZP;
    constexpr int MOUT = zf_mout<N, EPT>();
    constexpr int MST = ((NZP < N ? NZP : N) - 1) / TPF + 1;  // slots holding some k < min(NZP, N)
    zfft<N, EPT, -1, SY, TAB>(z, t, lds, base, wz);
    C o1[MST], o2[MST];
    if constexpr (PAIR) {
        const int src = (lane - t) + ((TPF - t) & (TPF - 1));
#pragma unroll
        for (int m = 0; m < MOUT; ++m) {
            C zp;  // Z[N - k], k = t + TPF m
            zp.x = __shfl(z[EPT - 1 - m].x, src);
            zp.y = __shfl(z[EPT - 1 - m].y, src);
            if (t == 0) zp = z[(EPT - m) % EPT];
            const C zk = z[m];
            o1[m] = cmk<C>(R(0.5) * (zk.x + zp.x), R(0.5) * (zk.y - zp.y));
            o2[m] = cmk<C>(R(0.5) * (zk.y + zp.y), R(0.5) * (zp.x - zk.x));
        }
    } else {
#pragma unroll
        for (int m = 0; m < MOUT; ++m) o1[m] = z[m];
    }
#pragma unroll
    for (int m = 0; m < MST; ++m) {
        const int k = t + TPF * m;
        if (k < NZP) {
            const bool in = (m < MOUT) && k <= kmax;
            if (live) out1[k] = in ? o1[m < MOUT ? m : 0] : cmk<C>(0, 0);
            if constexpr (PAIR) {
                if (live2) out2[k] = in ? o2[m < MOUT ? m : 0] : cmk<C>(0, 0);
            }
        }
    }
    if constexpr (NZP > N) {
        for (int k = N + t; k < NZP; k += TPF) {
            if (live) out1[k] = cmk<C>(0, 0);
            if constexpr (PAIR) {
                if (live2) out2[k] = cmk<C>(0, 0);
            }
        }
    }
}

// ZPAIR: when the line groups of a box come in pairs ((N*N) % (2*LPW) == 0) a wavefront takes two
// consecutive groups per iteration, and the real P_z of line l (first group) and of line l + LPW
// (second group) share one forward transform, P_z(l) + i P_z(l + LPW), separated like P_x + i P_y.
// Four complex transforms per line instead of five.  The partner of a line is fixed by the line
// alone (box offsets are multiples of N*N, hence of 2*LPW), so a box's bits do not depend on the
// grid or on the other boxes of a batch.
template <int N, int EPT> constexpr bool zfull_pairs() { return (N * N) % (2 * (64 / (N / EPT))) == 0; }
template <typename T, int N, int EPT, int WPB, int OCC, bool TAB, bool ZPAIR = true>
__global__ void __launch_bounds__(64 * WPB, OCC) passz_full_kernel(PZArgs a) {
    using C = cpx_t<T>;
    constexpr int TPF = N / EPT, LPW = 64 / TPF;
    static_assert(N % EPT == 0 && TPF <= 64 && 64 % TPF == 0, "a line's threads must divide a wavefront");
    constexpr bool PZ = ZPAIR && zfull_pairs<N, EPT>();
    constexpr int GPI = PZ ? 2 : 1;  // line groups per iteration
    constexpr int SY = SYNC_WAVE;
    constexpr int MP = fft_lds_len(N, EPT, true);
    __shared__ C lds[WPB * LPW * MP];
    // the wavefront index is uniform across the wavefront: held in a scalar register, so with one
    // line per wavefront (LPW = 1) every line address is scalar and the loads and stores take a
    // scalar base plus a per-thread offset that does not change from line to line
    const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int sub = lane / TPF, t = lane % TPF;
    C wz[zntw2<N, EPT, TAB>()];
    if constexpr (TAB) fft_load_twiddle_table<N, EPT>(wz, t, (const C*)a.twN);
    else fft_load_base_twiddles<N, EPT>(wz, t, (const C*)a.twN);
    const int base = (wv * LPW + sub) * MP;
    constexpr int NZP = Ext<N>::NZP;
    const long long nlines = (long long)N * N;
    const long long total = nlines * a.nbox;
    const long long ngroups = (total + LPW - 1) / LPW;
    const long long niter = PZ ? ngroups / 2 : ngroups;  // PZ: total is a multiple of 2 LPW
    const long long fstride = nlines * NZP;
    const C* in0 = (const C*)a.in;
    C* out0 = (C*)a.out;
    auto line_off = [&](long long gl, long long stride) {
        const long long b = gl / nlines;
        return b * stride + (gl - b * nlines) * NZP;
    };
    // running max of |u_x| + |u_y| + |u_z| in the transform precision (NaN propagates)
    using R = decltype(C::x);
    R mx = R(0);
    int mx_box = -1;
    auto flush_max = [&]() {
        if (a.track_max && mx_box >= 0) {
            double v = (double)mx * a.scale;
            for (int off = 32; off > 0; off >>= 1) v = nanmax(v, __shfl_xor(v, off));
            if (lane == 0) atomicMax(&a.sc[mx_box].maxbits, (unsigned long long)__double_as_longlong(v));
        }
        mx = R(0);
    };

    for (long long h = (long long)blockIdx.x * WPB + wv; h < niter; h += (long long)gridDim.x * WPB) {
        R pz0[PZ ? EPT : 1];  // P_z of the iteration's first line (PZ)
        C* out_first = nullptr;
#pragma unroll 1
        for (int j = 0; j < GPI; ++j) {
            const long long g = h * GPI + j;
            long long gl = g * LPW + sub;
            const bool live = gl < total;
            if (!live) gl = total - 1;  // idle lanes recompute the last line but never store
            // boxes of the group's first and last line (uniform); they differ when N^2 is not a
            // multiple of the lines per wavefront (N = 12: 144 lines, 32 per wavefront)
            const int box = (int)((g * LPW) / nlines);
            const int box_last = (int)(((g + 1) * LPW - 1 < total ? (g + 1) * LPW - 1 : total - 1) / nlines);
            if (box != mx_box) {
                flush_max();
                mx_box = box;
            }
            const C* in = in0 + line_off(gl, a.ibox);
            C* out = out0 + line_off(gl, a.obox);
            C a1[EPT], a2[EPT], a3[EPT], cu[EPT], cw[EPT];
            // (u_y, w_z): P_x = u_y w_z
            zf_load<T, N, EPT>(cu, in + 1 * fstride, t);  // u_y
            zf_load<T, N, EPT>(cw, in + 5 * fstride, t);  // w_z
            zf_pack<T, N, EPT>(a1, cu, cw, t);
            zf_load<T, N, EPT>(cu, in + 2 * fstride, t);  // u_z
            zf_load<T, N, EPT>(cw, in + 4 * fstride, t);  // w_y
            zfft<N, EPT, +1, SY, TAB>(a1, t, lds, base, wz);
            // (u_z, w_y): P_x -= u_z w_y
            zf_pack<T, N, EPT>(a2, cu, cw, t);
            zf_load<T, N, EPT>(cu, in + 0 * fstride, t);  // u_x
            zf_load<T, N, EPT>(cw, in + 3 * fstride, t);  // w_x
            zfft<N, EPT, +1, SY, TAB>(a2, t, lds, base, wz);
            // (u_x, w_x)
            zf_pack<T, N, EPT>(a3, cu, cw, t);
            zfft<N, EPT, +1, SY, TAB>(a3, t, lds, base, wz);
            // a1 = (u_y, w_z), a2 = (u_z, w_y), a3 = (u_x, w_x) at the grid points t + TPF m
            C pq[EPT];
            R pz[EPT];
#pragma unroll
            for (int m = 0; m < EPT; ++m) {
                const auto uy = a1[m].x, wzv = a1[m].y, uz = a2[m].x, wy = a2[m].y, ux = a3[m].x, wx = a3[m].y;
                pq[m] = cmk<C>(uy * wzv - uz * wy, uz * wx - ux * wzv);  // P_x + i P_y
                pz[m] = ux * wy - uy * wx;
            }
            if (a.track_max) {
                R ml = R(0);
#pragma unroll
                for (int m = 0; m < EPT; ++m) {
                    const R v = fabs(a3[m].x) + fabs(a1[m].x) + fabs(a2[m].x);
                    ml = (v >= ml || v != v) ? v : ml;
                }
                if (!live) ml = R(0);
                if (box == box_last) {
                    mx = (ml >= mx || ml != ml) ? ml : mx;
                } else {
                    // the group straddles boxes box and box + 1: reduce each box's part now
                    const bool mine = (int)(gl / nlines) == box;
                    double va = mine ? (double)ml * a.scale : 0.0, vb = mine ? 0.0 : (double)ml * a.scale;
                    for (int off = 32; off > 0; off >>= 1) {
                        va = nanmax(va, __shfl_xor(va, off));
                        vb = nanmax(vb, __shfl_xor(vb, off));
                    }
                    if (lane == 0) {
                        atomicMax(&a.sc[box].maxbits, (unsigned long long)__double_as_longlong(va));
                        atomicMax(&a.sc[box_last].maxbits, (unsigned long long)__double_as_longlong(vb));
                    }
                }
            }
            zf_forward_store<T, N, EPT, SY, TAB, true>(pq, out + 0 * fstride, out + 1 * fstride, t, lane, lds, base, wz,
                                                       live, live);
            C zz[EPT];
            if constexpr (PZ) {
                if (j == 0) {
#pragma unroll
                    for (int m = 0; m < EPT; ++m) pz0[m] = pz[m];
                    out_first = out;
                    continue;
                }
                // every line of a paired iteration is live (total is a multiple of 2 LPW)
#pragma unroll
                for (int m = 0; m < EPT; ++m) zz[m] = cmk<C>(pz0[m], pz[m]);
                zf_forward_store<T, N, EPT, SY, TAB, true>(zz, out_first + 2 * fstride, out + 2 * fstride, t, lane, lds,
                                                           base, wz, true, true);
            } else {
                // P_z alone (real input)
#pragma unroll
                for (int m = 0; m < EPT; ++m) zz[m] = cmk<C>(pz[m], R(0));
                zf_forward_store<T, N, EPT, SY, TAB, false>(zz, out + 2 * fstride, nullptr, t, lane, lds, base, wz, live,
                                                            false);
            }
        }
    }
    flush_max();
}

// ------------------------------------------------------------------------------------------
// Pass 5: forward x transform + epilogue.
//  EP_STAGE: NL = r2c(...)/N^6, Leray projection (k=0 -> 0), 2/3 mask, forcing, then
//            stage < 4: K_s = NL;  stage 4: u_hat = gamma u_hat + sum delta_i K_i
//  EP_STORE: u_hat = scale * X masked (initial conditions from real space)
// grid (NZP/W, NK); block W*TPF
// ------------------------------------------------------------------------------------------
enum EpMode { EP_STAGE = 0, EP_STORE = 1 };

// batches: box = blockIdx.z; per-box element strides sbox (state/K) and ibox (A4)
struct P5Args {
    const void* in;   // A4
    double2* uh;      // state at the start of the step
    double2* unew;    // Dormand-Prince final stage: candidate state (never uh), or null (in place)
    StageRow row;     // K_1..K_6 (bulk precision) and the stage's tableau row (Dormand-Prince)
    void* kout;       // K_s output (stage < nstage)
    double* errpart;  // Dormand-Prince final stage: per-block error sums [box][slot][3], or null
    const double2* extra;  // optional term added to every stage's nonlinear term (state layout)
    const void* twN;
    const DevScalars* sc;
    const double* fslots;
    double kf2, eps_in, norm;
    long long sbox, ibox;
    int visc_pow, scheme, stage, nstage, force, nslots;
    int kmax, NK, NZP;
};

// Dormand-Prince 5(4) weights (Solver/src/solver.c:41-42) and error weights B_i - Bs_i
constexpr double DPB1 = 35.0 / 384.0, DPB3 = 500.0 / 1113.0, DPB4 = 125.0 / 192.0, DPB5 = -2187.0 / 6784.0,
                 DPB6 = 11.0 / 84.0;
constexpr double DPE1 = DPB1 - 5179.0 / 57600.0, DPE3 = DPB3 - 7571.0 / 16695.0, DPE4 = DPB4 - 393.0 / 640.0,
                 DPE5 = DPB5 - (-92097.0 / 339200.0), DPE6 = DPB6 - 187.0 / 2100.0, DPE7 = -1.0 / 40.0;
// error tolerances (Solver/src/data_types.h:123-124)
constexpr double DP_ATOL = 1e-7, DP_RTOL = 1e-7;

// Reads of the final RK4 update of slot m (state and K1..K3 of the three components), when the slot
// is stored.
template <typename T, int N, int EPT, bool O32>
__device__ __forceinline__ void p5_fetch(int m, int t, int iy, int kz, int kmax, int NK, int NZP, long long plane,
                                         const double2* uh, const cpx_t<T>* K1, const cpx_t<T>* K2, const cpx_t<T>* K3,
                                         double2 (&u0)[3], cpx_t<T> (&q1)[3], cpx_t<T> (&q2)[3], cpx_t<T> (&q3)[3]) {
    const int kx = k_of_pos(t + (N / EPT) * m, N);
    if (kx < -kmax || kx > kmax || kz >= NZP) return;
    const long long j = sidx(cidx_of_k(kx, NK), iy, kz, NK, NZP);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        u0[c] = at<O32>(uh + c * plane, j);
        q1[c] = at<O32>(K1 + c * plane, j);
        q2[c] = at<O32>(K2 + c * plane, j);
        q3[c] = at<O32>(K3 + c * plane, j);
    }
}

// Epilogue of one (box, row iy, column kz) after the forward x transforms: r holds the three fields
// at the x positions t + TPF m; es accumulates this thread's Dormand-Prince error sums.  SCH fixes
// the scheme at compile time (RK4CN or IFRK4; -1: the runtime a.scheme), so that an RK4-CN
// instance carries no exponential.
template <typename T, int N, int EPT, int MODE, int FAM, int SCH = -1>
__device__ __forceinline__ void p5_epilogue(const P5Args& a, const cpx_t<T> (&r)[3][EPT], int box, int iy, int kz, int t,
                                            double (&es)[3]) {
    using C = cpx_t<T>;
    constexpr int TPF = N / EPT;
    constexpr bool O32 = N <= 512;  // every component plane below 4 GiB
    const int scheme = (SCH >= 0) ? SCH : a.scheme;
    const DevScalars& sc = a.sc[box];
    double2* __restrict__ uh = a.uh + box * a.sbox;
    double2* __restrict__ unew = a.unew ? a.unew + box * a.sbox : nullptr;  // null: update in place
    const C* K1 = (const C*)a.row.k[0] + box * a.sbox;
    const C* K2 = (const C*)a.row.k[1] + box * a.sbox;
    const C* K3 = (const C*)a.row.k[2] + box * a.sbox;
    C* ko = a.kout ? (C*)a.kout + box * a.sbox : nullptr;
    const double2* extra = a.extra ? a.extra + box * a.sbox : nullptr;
    const double* fslots = a.fslots + box * a.NK;
    const int kmax = a.kmax, NK = a.NK, NZP = a.NZP;
    const int ky = k_of_cidx(iy, kmax, NK);
    const long long plane = (long long)NK * NK * NZP;

    double ffac = 0.0;
    if (MODE == EP_STAGE && a.force) {
        // forced-band energy of the stage input, summed over the pass-1 slots in fixed order
        double s = 0.0;
        int kfi = (int)floor(sqrt(a.kf2));
        if (kfi > kmax) kfi = kmax;
        for (int q = -kfi; q <= kfi; ++q) s += fslots[cidx_of_k(q, NK)];
        const double N3 = (double)N * N * N;
        const double Ef = 0.5 * s / (N3 * N3);
        ffac = (Ef > 0.0) ? a.eps_in / (2.0 * Ef) : 0.0;
    }
    const double dt = sc.dt, nu = sc.nu;
    const long long kcut2 = (long long)kmax * kmax;

    // final RK4 stage with FP32 slopes: the twelve reads of slot m's update (state and K1..K3 of the
    // three components) are issued together before slot m's term is computed, instead of four at a
    // time behind the previous component's store (in issue order a read behind a store waits for
    // it); FP64 keeps the per-component reads (its batch does not fit the registers)
    constexpr bool BATCH = MODE == EP_STAGE && FAM == 0 && sizeof(C) == 8;
    const bool fin = BATCH && a.stage >= a.nstage;
    double2 u0b[3];
    C q1b[3], q2b[3], q3b[3];

#pragma unroll
    for (int m = 0; m < EPT; ++m) {
        if (fin) p5_fetch<T, N, EPT, O32>(m, t, iy, kz, kmax, NK, NZP, plane, uh, K1, K2, K3, u0b, q1b, q2b, q3b);
        const int p = t + TPF * m;
        const int kx = k_of_pos(p, N);
        if (kx < -kmax || kx > kmax || kz >= NZP) continue;
        const long long idx = sidx(cidx_of_k(kx, NK), iy, kz, NK, NZP);
        const long long k2 = (long long)kx * kx + (long long)ky * ky + (long long)kz * kz;
        const bool keep = k2 <= kcut2 && kz <= kmax;
        double2 nl[3];
        if (MODE == EP_STORE) {
#pragma unroll
            for (int c = 0; c < 3; ++c)
                nl[c] = keep ? make_double2((double)r[c][m].x * a.norm, (double)r[c][m].y * a.norm) : make_double2(0.0, 0.0);
#pragma unroll
            for (int c = 0; c < 3; ++c) at<O32>(uh + c * plane, idx) = nl[c];
            continue;
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) nl[c] = make_double2((double)r[c][m].x * a.norm, (double)r[c][m].y * a.norm);
        if (k2 != 0 && keep) {
            const double kinv = 1.0 / (double)k2;
            const double kxd = kx, kyd = ky, kzd = kz;
            const double dre = kxd * nl[0].x + kyd * nl[1].x + kzd * nl[2].x;
            const double dim = kxd * nl[0].y + kyd * nl[1].y + kzd * nl[2].y;
            nl[0].x -= kxd * kinv * dre; nl[0].y -= kxd * kinv * dim;
            nl[1].x -= kyd * kinv * dre; nl[1].y -= kyd * kinv * dim;
            nl[2].x -= kzd * kinv * dre; nl[2].y -= kzd * kinv * dim;
        } else {
#pragma unroll
            for (int c = 0; c < 3; ++c) nl[c] = make_double2(0.0, 0.0);
        }
        const double D = visc_rate(nu, a.visc_pow, k2);
        if (a.force && keep && k2 > 0 && (double)k2 <= a.kf2) {
            if constexpr (FAM == 1) {
                double2 us[3];
#pragma unroll
                for (int c = 0; c < 3; ++c) us[c] = at<O32>(uh + c * plane, idx);
                add_slopes<T>(us, a.row, a.stage - 1, box * a.sbox, idx, plane, dt);
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    nl[c].x += ffac * us[c].x;
                    nl[c].y += ffac * us[c].y;
                }
            } else {
                double al, be;
                stage_in_coef(scheme, a.stage, D, dt, al, be);
                const C* kp = (a.stage == 2) ? K1 : (a.stage == 3) ? K2 : K3;
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    double2 us = at<O32>(uh + c * plane, idx);
                    if (a.stage > 1) {
                        double2 kv = to_d2<T>(at<O32>(kp + c * plane, idx));
                        us = make_double2(al * us.x + be * kv.x, al * us.y + be * kv.y);
                    }
                    nl[c].x += ffac * us.x;
                    nl[c].y += ffac * us.y;
                }
            }
        }
        if (extra && keep) {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const double2 e = at<O32>(extra + c * plane, idx);
                nl[c].x += e.x;
                nl[c].y += e.y;
            }
        }
        if (a.stage < a.nstage) {
#pragma unroll
            for (int c = 0; c < 3; ++c) at<O32>(ko + c * plane, idx) = from_d2<T>(nl[c]);
        } else if constexpr (FAM == 1) {
            // u (2-D)/(2+D) + 2dt/(2+D) sum_i B_i K_i (Solver/src/solver.c:336-349); Dormand-Prince:
            // the candidate goes to unew and e = 2dt/(2+D) sum_i (B_i - Bs_i) K_i is measured
            // against atol + max(|u|, |u5|) rtol per component (:356-384, estimate corrected)
            const double Dd = D * dt;
            const double g = (2.0 - Dd) / (2.0 + Dd), h = 2.0 * dt / (2.0 + Dd);
            const C* K4 = (const C*)a.row.k[3] + box * a.sbox;
            const C* K5 = (const C*)a.row.k[4] + box * a.sbox;
            const bool dp = a.nstage == 7;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const double2 u0 = at<O32>(uh + c * plane, idx);
                const double2 q1 = to_d2<T>(at<O32>(K1 + c * plane, idx));
                const double2 q3 = to_d2<T>(at<O32>(K3 + c * plane, idx));
                const double2 q4 = to_d2<T>(at<O32>(K4 + c * plane, idx));
                const double2 q5 = to_d2<T>(at<O32>(K5 + c * plane, idx));
                double2 q6 = nl[c];
                if (dp) q6 = to_d2<T>(at<O32>((const C*)a.row.k[5] + box * a.sbox + c * plane, idx));
                double2 un;
                un.x = u0.x * g + h * (DPB1 * q1.x + DPB3 * q3.x + DPB4 * q4.x + DPB5 * q5.x + DPB6 * q6.x);
                un.y = u0.y * g + h * (DPB1 * q1.y + DPB3 * q3.y + DPB4 * q4.y + DPB5 * q5.y + DPB6 * q6.y);
                if (!keep) un = make_double2(0.0, 0.0);
                if (dp) {
                    double ex = DPE1 * q1.x + DPE3 * q3.x + DPE4 * q4.x + DPE5 * q5.x + DPE6 * q6.x + DPE7 * nl[c].x;
                    double ey = DPE1 * q1.y + DPE3 * q3.y + DPE4 * q4.y + DPE5 * q5.y + DPE6 * q6.y + DPE7 * nl[c].y;
                    ex *= h;
                    ey *= h;
                    const double a0 = sqrt(u0.x * u0.x + u0.y * u0.y), a1 = sqrt(un.x * un.x + un.y * un.y);
                    const double den = DP_ATOL + (a0 > a1 ? a0 : a1) * DP_RTOL;
                    if (keep) es[c] += (ex * ex + ey * ey) / (den * den);
                    at<O32>(unew + c * plane, idx) = un;
                } else {
                    at<O32>(uh + c * plane, idx) = un;
                }
            }
        } else {
            constexpr double B1 = 1.0 / 6.0, B2 = 1.0 / 3.0, B3 = 1.0 / 3.0, B4 = 1.0 / 6.0;
            double g, h = 0.0, e2 = 0.0;
            if (scheme == RK4CN) {
                const double Dd = D * dt;
                g = (2.0 - Dd) / (2.0 + Dd);
                h = 2.0 * dt / (2.0 + Dd);
            } else {
                g = exp(-D * dt);
                e2 = exp(-D * dt * 0.5);
            }
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                double2 u0, q1, q2, q3;
                if constexpr (BATCH) {
                    u0 = u0b[c];
                    q1 = to_d2<T>(q1b[c]);
                    q2 = to_d2<T>(q2b[c]);
                    q3 = to_d2<T>(q3b[c]);
                } else {
                    u0 = at<O32>(uh + c * plane, idx);
                    q1 = to_d2<T>(at<O32>(K1 + c * plane, idx));
                    q2 = to_d2<T>(at<O32>(K2 + c * plane, idx));
                    q3 = to_d2<T>(at<O32>(K3 + c * plane, idx));
                }
                double2 un;
                if (scheme == RK4CN) {
                    // u (2-D)/(2+D) + 2dt/(2+D) (K1/6 + K2/3 + K3/3 + K4/6), Solver/src/solver.c:601-603
                    un.x = u0.x * g + h * (B1 * q1.x + B2 * q2.x + B3 * q3.x + B4 * nl[c].x);
                    un.y = u0.y * g + h * (B1 * q1.y + B2 * q2.y + B3 * q3.y + B4 * nl[c].y);
                } else {
                    // E u + dt/6 (E K1 + 2 E2 (K2 + K3) + K4)
                    un.x = g * u0.x + dt / 6.0 * (g * q1.x + 2.0 * e2 * (q2.x + q3.x) + nl[c].x);
                    un.y = g * u0.y + dt / 6.0 * (g * q1.y + 2.0 * e2 * (q2.y + q3.y) + nl[c].y);
                }
                if (!keep) un = make_double2(0.0, 0.0);
                at<O32>(uh + c * plane, idx) = un;
            }
        }
    }
}

template <typename T, int N, int EPT, int W, int MODE, int FAM>
__global__ void __launch_bounds__(N / EPT * W) pass5_kernel(P5Args a) {
    using C = cpx_t<T>;
    constexpr int TPF = N / EPT;
    constexpr bool PAIR = W * sizeof(C) < 128;
    constexpr bool O32 = N <= 512;  // every component plane below 4 GiB
    __shared__ C lds[N * W];
    int tile, iy;
    if (!tile_of_block<PAIR>(a.NZP / W, a.NK, tile, iy)) return;
    const int box = blockIdx.z;
    const int tid = threadIdx.x, w = tid % W, t = tid / W;
    const int kz = tile * W + w;
    const int NK = a.NK, NZP = a.NZP;
    const long long iplane = (long long)N * xplane(NK, NZP);
    const C* in = (const C*)a.in + box * a.ibox;
    const C* tw = (const C*)a.twN;

    C r[3][EPT];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
#pragma unroll
        for (int m = 0; m < EPT; ++m) r[c][m] = at<O32>(in + c * iplane, xidx<N>(t + TPF * m, iy, kz, NK, NZP));
        fft<N, EPT, -1>(r[c], t, lds, w, W, tw);
    }
    double es[3] = {0.0, 0.0, 0.0};  // Dormand-Prince error sums of this thread
    p5_epilogue<T, N, EPT, MODE, FAM>(a, r, box, iy, kz, t, es);
    if (FAM == 1 && MODE == EP_STAGE && a.errpart) {
        // per-block error sums, reduced in a fixed order by dp_finalize_kernel
        __syncthreads();  // the transforms are done with lds
        double* red = (double*)lds;
        double* dst = a.errpart + ((long long)box * a.nslots + (long long)iy * (NZP / W) + tile) * 3;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double v = block_sum(es[c], red);
            if (tid == 0) dst[c] = v;
        }
    }
}


// Pass 5 with its field reads pipelined (RK4 family, stage epilogue, one 128-byte segment per x
// line: W * sizeof(C) == 128).  Persistent blocks walk the (box, row iy, column tile) items; each
// field's N x W input goes from HBM straight into an LDS staging buffer through the global->LDS
// DMA (64 lanes x 16 bytes per instruction, eight x lines) while the block transforms the field
// before, so the reads of one field overlap the transform of the previous one.  (pass5_kernel
// holds one such block per CU, which issues all its reads, then transforms with none in flight.)
// grid: at most one block per CU (2 x N*W elements of LDS); items = boxes * NK * (NZP / W).
// SCH: the scheme (RK4CN or IFRK4), fixed at compile time.
template <typename T, int N, int EPT, int W, int SCH>
__global__ void __launch_bounds__(N / EPT * W) pass5p_kernel(P5Args a, int ntiles, int nitems) {
    using C = cpx_t<T>;
    constexpr int TPF = N / EPT, NT = TPF * W, NWAVE = NT / 64;
    static_assert(W * sizeof(C) == 128, "one 128-byte segment per x line");
    static_assert(NT % 64 == 0, "whole wavefronts");
    constexpr int ROWS = 8;                // x lines per DMA instruction (64 lanes x 16 bytes)
    static_assert(N % (ROWS * NWAVE) == 0, "whole DMA instructions per wavefront");
    constexpr int INST = N / (ROWS * NWAVE);  // DMA instructions per wavefront and field
    constexpr int EPL = 16 / (int)sizeof(C);  // complex elements per lane and instruction
    __shared__ C lds[N * W];  // transform exchange
    __shared__ C stg[N * W];  // staged input of one field, [x][W]
    __shared__ C twl[N];      // twiddle table: transform reads from LDS keep the global counter free
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int NK = a.NK, NZP = a.NZP;
    const long long xpl = xplane(NK, NZP), iplane = (long long)N * xpl;
    for (int e = threadIdx.x; e < N; e += NT) twl[e] = ((const C*)a.twN)[e];
    __syncthreads();
    // DMA of field c of an item: lane l of instruction i of wavefront wv copies 16 bytes of x line
    // (wv INST + i) ROWS + l / 8, columns tile W + (l % 8) EPL ..., to the same place in stg
    auto issue = [&](int item, int c, int lane) {
        const int tile = item % ntiles, rest = item / ntiles, iy = rest % NK, box = rest / NK;
        const C* src = (const C*)a.in + box * a.ibox + c * iplane + (long long)iy * NZP + tile * W + (lane % 8) * EPL;
#pragma unroll
        for (int i = 0; i < INST; ++i) {
            const int row0 = (wv * INST + i) * ROWS;
            const C* g = src + (long long)(row0 + lane / 8) * xpl;
            __builtin_amdgcn_global_load_lds((const void*)g, (__attribute__((address_space(3))) void*)(stg + row0 * W), 16, 0,
                                             0);
        }
    };
    int item = blockIdx.x;
    if (item < nitems) issue(item, 0, threadIdx.x & 63);
    for (; item < nitems; item += gridDim.x) {
        // the thread's coordinates, opaque to the compiler on every item: what derives from them
        // (transform and epilogue addresses) is recomputed per item instead of held in registers
        // across the loop
        int tid = threadIdx.x;
        asm volatile("" : "+v"(tid));
        const int w = tid % W, t = tid / W, lane = tid & 63;
        const int tile = item % ntiles, rest = item / ntiles, iy = rest % NK, box = rest / NK;
        const int kz = tile * W + w;
        C r[3][EPT];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            // this wavefront's copies (and stores) are done, then every wavefront's
            asm volatile("s_waitcnt vmcnt(0) lgkmcnt(0)\n\ts_barrier" ::: "memory");
#pragma unroll
            for (int m = 0; m < EPT; ++m) r[c][m] = stg[(t + TPF * m) * W + w];
            lds_sync<SYNC_LDS>();  // stg read by every thread before it is refilled
            const int nitem = (c < 2) ? item : item + (int)gridDim.x;
            if (nitem < nitems) issue(nitem, (c + 1) % 3, lane);
            fft<N, EPT, -1, SYNC_LDS>(r[c], t, lds, w, W, twl);
        }
        double es[3] = {0.0, 0.0, 0.0};
        p5_epilogue<T, N, EPT, EP_STAGE, 0, SCH>(a, r, box, iy, kz, t, es);
    }
}

// ------------------------------------------------------------------------------------------
// Step bookkeeping (single-thread kernels; no memsets so the step can be graph-captured)
// ------------------------------------------------------------------------------------------
__global__ void step_begin_kernel(DevScalars* sc, int nbox) {
    for (int b = threadIdx.x; b < nbox; b += blockDim.x) sc[b].maxbits = 0ull;
}
__global__ void cfl_kernel(DevScalars* sc, int nbox) {
    for (int b = threadIdx.x; b < nbox; b += blockDim.x) {
        const double mx = __longlong_as_double((long long)sc[b].maxbits);
        if (sc[b].cfl > 0.0 && mx > 0.0) sc[b].dt = sc[b].cfl * sc[b].dx / mx;
    }
}
__global__ void step_end_kernel(DevScalars* sc, int nbox) {
    for (int b = threadIdx.x; b < nbox; b += blockDim.x) {
        sc[b].t += sc[b].dt;
        sc[b].steps += 1;
        sc[b].max_last = __longlong_as_double((long long)sc[b].maxbits);
    }
}

// Dormand-Prince step control (Solver/src/solver.c:134-153, constants Solver/src/data_types.h:125-127):
// err = max_c sqrt(sum_c / (Nx Ny Nzf)); accepted when err <= 1 (then t += dt); in every case
// dt <- dt min(1.5, max(0.01, 0.8 err^(-1/5))), with the reference's max/min, which pass a NaN on.
// One block per box; the per-block sums are added in a fixed order.
__device__ __forceinline__ double dp_max(double a, double b) { return a > b ? a : b; }
__device__ __forceinline__ double dp_min(double a, double b) { return a < b ? a : b; }
__global__ void __launch_bounds__(256) dp_finalize_kernel(DevScalars* sc, const double* errpart, int nslots, double count) {
    __shared__ double red[256];
    const int box = blockIdx.x;
    const double* part = errpart + (long long)box * nslots * 3;
    double s[3] = {0.0, 0.0, 0.0};
    for (int j = threadIdx.x; j < nslots; j += 256) {
#pragma unroll
        for (int c = 0; c < 3; ++c) s[c] += part[3 * j + c];
    }
    double err = 0.0;
#pragma unroll
    for (int c = 0; c < 3; ++c) err = nanmax(err, sqrt(block_sum(s[c], red) / count));
    if (threadIdx.x == 0) {
        DevScalars& d = sc[box];
        const double h = d.dt;
        d.dperr = err;
        d.dpacc = err <= 1.0 ? 1 : 0;
        if (err <= 1.0) {
            d.t += h;
            d.steps += 1;
            d.max_last = __longlong_as_double((long long)d.maxbits);
        }
        d.dt = h * dp_min(1.5, dp_max(0.01, 0.8 * pow(1.0 / err, 0.2)));
    }
}

// ------------------------------------------------------------------------------------------
// Layout conversion: reference AoS [N][N][Nzf][3] (interleaved complex) <-> compact state
// ------------------------------------------------------------------------------------------
__global__ void aos_to_state_kernel(const double2* __restrict__ aos, double2* __restrict__ uh, int N, int kmax, int NK, int NZP) {
    const long long plane = (long long)NK * NK * NZP;
    const long long total = plane;
    const long long kcut2 = (long long)kmax * kmax;
    for (long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x; n < total; n += (long long)gridDim.x * blockDim.x) {
        const int kz = (int)(n % NZP);
        const int ix = (int)((n / NZP) % NK);
        const int iy = (int)(n / ((long long)NZP * NK));
        const int kx = k_of_cidx(ix, kmax, NK), ky = k_of_cidx(iy, kmax, NK);
        const long long k2 = (long long)kx * kx + (long long)ky * ky + (long long)kz * kz;
        const bool keep = kz <= kmax && k2 <= kcut2;
        const int Nzf = N / 2 + 1;
        const long long src = (((long long)pos_of_k(kx, N) * N + pos_of_k(ky, N)) * Nzf + kz) * 3;
#pragma unroll
        for (int c = 0; c < 3; ++c) uh[c * plane + n] = keep ? aos[src + c] : make_double2(0.0, 0.0);
    }
}

template <typename CT>
__global__ void compact_to_aos_kernel(const CT* __restrict__ src, double2* __restrict__ aos, int N, int kmax, int NK, int NZP) {
    const int Nzf = N / 2 + 1;
    const long long total = (long long)N * N * Nzf;
    const long long plane = (long long)NK * NK * NZP;
    for (long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x; n < total; n += (long long)gridDim.x * blockDim.x) {
        const int kz = (int)(n % Nzf);
        const int j = (int)((n / Nzf) % N);
        const int i = (int)(n / ((long long)Nzf * N));
        const int kx = k_of_pos(i, N), ky = k_of_pos(j, N);
        const bool in = kz <= kmax && kx >= -kmax && kx <= kmax && ky >= -kmax && ky <= kmax;
        const long long s = sidx(cidx_of_k(kx, NK), cidx_of_k(ky, NK), kz, NK, NZP);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            double2 v = make_double2(0.0, 0.0);
            if (in) { CT q = src[c * plane + s]; v = make_double2((double)q.x, (double)q.y); }
            aos[n * 3 + c] = v;
        }
    }
}

// ------------------------------------------------------------------------------------------
// Diagnostics: ComputeSystemMeasurables (Solver/src/solver.c:1142-1276), one block per shell b
// (modes with round(|k|) = b, i.e. k^2 in [b^2-b+1, b^2+b]); deterministic fixed-order sums.
// part[b*7 + {0..6}] = {E, Ens, Pal, Hel, Diss, E_b, Ens_b} before the global scale factors
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ long long isqrt_ll(long long v) {
    if (v <= 0) return 0;
    long long s = (long long)sqrt((double)v);
    while (s * s > v) --s;
    while ((s + 1) * (s + 1) <= v) ++s;
    return s;
}

// grid (n_spect, nbox), block <= 1024 threads (one shell per block: the outer shells hold up to
// ~4 pi kmax^2 modes, so wide blocks keep the largest shell from setting the time);
// part[(box * n_spect + b) * 7 + j]
__global__ void measures_kernel(const double2* __restrict__ uh0, double* __restrict__ part0, int N, int kmax, int NK,
                                int NZP, const DevScalars* __restrict__ sc, int visc_pow, int quirk, long long sbox) {
    __shared__ double red[1024];
    const int b = blockIdx.x, box = blockIdx.y;
    const double2* __restrict__ uh = uh0 + box * sbox;
    double* __restrict__ part = part0 + (long long)box * gridDim.x * 7;
    const double nu = sc[box].nu;
    const long long lo = (b == 0) ? 0 : (long long)b * b - b + 1;
    const long long hi = (b == 0) ? 0 : (long long)b * b + b;
    const long long plane = (long long)NK * NK * NZP;
    const int Nzf = N / 2 + 1;
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int q = threadIdx.x; q < NK * NK; q += blockDim.x) {
        const int ix = q / NK, iy = q % NK;
        const long long kx = k_of_cidx(ix, kmax, NK), ky = k_of_cidx(iy, kmax, NK);
        const long long kh2 = kx * kx + ky * ky;
        if (kh2 > hi) continue;
        const long long zlo_sq = lo - kh2;
        long long zlo = (zlo_sq <= 0) ? 0 : isqrt_ll(zlo_sq - 1) + 1;  // smallest kz with kz^2 >= lo - kh2
        long long zhi = isqrt_ll(hi - kh2);
        if (zhi > kmax) zhi = kmax;
        for (long long kz = zlo; kz <= zhi; ++kz) {
            const long long idx = sidx(ix, iy, kz, NK, NZP);
            const double kxd = (double)kx, kyd = (double)ky, kzd = (double)kz;
            double2 u[3], w[3], c[3];
            for (int d = 0; d < 3; ++d) u[d] = uh[d * plane + idx];
            {
                double ar = kyd * u[2].x - kzd * u[1].x, ai = kyd * u[2].y - kzd * u[1].y;
                double br = kzd * u[0].x - kxd * u[2].x, bi = kzd * u[0].y - kxd * u[2].y;
                double cr = kxd * u[1].x - kyd * u[0].x, ci = kxd * u[1].y - kyd * u[0].y;
                w[0] = make_double2(-ai, ar); w[1] = make_double2(-bi, br); w[2] = make_double2(-ci, cr);
            }
            {
                double ar = kyd * w[2].x - kzd * w[1].x, ai = kyd * w[2].y - kzd * w[1].y;
                double br = kzd * w[0].x - kxd * w[2].x, bi = kzd * w[0].y - kxd * w[2].y;
                double cr = kxd * w[1].x - kyd * w[0].x, ci = kxd * w[1].y - kyd * w[0].y;
                c[0] = make_double2(-ai, ar); c[1] = make_double2(-bi, br); c[2] = make_double2(-ci, cr);
            }
            double au[3], aw[3], ac[3];
            for (int d = 0; d < 3; ++d) {
                au[d] = u[d].x * u[d].x + u[d].y * u[d].y;
                aw[d] = w[d].x * w[d].x + w[d].y * w[d].y;
                ac[d] = c[d].x * c[d].x + c[d].y * c[d].y;
            }
            const long long k2 = kh2 + kz * kz;
            const double pre = visc_rate(nu, visc_pow, k2);
            double hel = 0.0;
            if (quirk) for (int d = 0; d < 3; ++d) hel += u[d].x * w[d].x - u[d].y * w[d].y;
            else for (int d = 0; d < 3; ++d) hel += u[d].x * w[d].x + u[d].y * w[d].y;
            const double U = au[0] + au[1] + au[2], Wv = aw[0] + aw[1] + aw[2], Cc = ac[0] + ac[1] + ac[2];
            const bool edge = (kz == 0 || kz == Nzf - 1);
            if (edge) {
                acc[4] += pre * U; acc[1] += Wv; acc[2] += Cc; acc[3] += hel; acc[0] += U;
            } else if (quirk) {
                acc[4] += 2.0 * pre * U;
                acc[1] += 2.0 * aw[0] + aw[1] + aw[2];
                acc[2] += 2.0 * ac[0] + ac[1] + ac[2];
                acc[3] += 2.0 * hel;
                acc[0] += 2.0 * au[0] + au[1] + au[2];
            } else {
                acc[4] += 2.0 * pre * U; acc[1] += 2.0 * Wv; acc[2] += 2.0 * Cc; acc[3] += 2.0 * hel; acc[0] += 2.0 * U;
            }
            const double wg = edge ? 1.0 : 2.0;
            acc[5] += wg * U;
            acc[6] += wg * Wv;
        }
    }
    for (int j = 0; j < 7; ++j) {
        double s = block_sum(acc[j], red);
        if (threadIdx.x == 0) part[b * 7 + j] = s;
    }
}

// ------------------------------------------------------------------------------------------
// Initial conditions
// ------------------------------------------------------------------------------------------
// real-space fields written as AoS [x][y][z][3]: kind 0 = Taylor-Green (Solver/src/solver.c:1553-1574),
// kind 1 = Shapiro with the divergence-free sin(Mz) (Solver/src/solver.c:1575-1600, D17 corrected)
__global__ void ic_real_kernel(double* __restrict__ u, int N, int kind, double nu, double t) {
    const long long total = (long long)N * N * N;
    for (long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x; n < total; n += (long long)gridDim.x * blockDim.x) {
        const int k = (int)(n % N), j = (int)((n / N) % N), i = (int)(n / ((long long)N * N));
        const double x = 2.0 * kPi * i / N, y = 2.0 * kPi * j / N, z = 2.0 * kPi * k / N;
        double ux, uy, uz;
        if (kind == 0) {
            ux = sin(x) * cos(y) * cos(z);
            uy = -cos(x) * sin(y) * cos(z);
            uz = 0.0;
        } else {
            const double A = 2.0, K = 2.0, L = 2.0, Mm = 2.0;
            const double lam = sqrt(K * K + L * L + Mm * Mm);
            const double dec = exp(-lam * lam * t * nu);
            ux = -A / (K * K + L * L) * (lam * L * cos(K * x) * sin(L * y) * sin(Mm * z) + Mm * K * sin(K * x) * cos(L * y) * cos(Mm * z)) * dec;
            uy = A / (K * K + L * L) * (lam * K * sin(K * x) * cos(L * y) * sin(Mm * z) - Mm * L * cos(K * x) * sin(L * y) * cos(Mm * z)) * dec;
            uz = A * cos(K * x) * cos(L * y) * sin(Mm * z) * dec;
        }
        u[n * 3 + 0] = ux;
        u[n * 3 + 1] = uy;
        u[n * 3 + 2] = uz;
    }
}

__device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
    for (int r = 0; r < 10; ++r) {
        const unsigned long long p0 = (unsigned long long)0xD2511F53u * c[0];
        const unsigned long long p1 = (unsigned long long)0xCD9E8D57u * c[2];
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        const uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}
__device__ __forceinline__ double u01(uint32_t a, uint32_t b) {
    unsigned long long v = ((unsigned long long)a << 21) ^ (unsigned long long)(b >> 11);
    v &= (1ull << 53) - 1;
    return ((double)v + 0.5) * (1.0 / 9007199254740992.0);
}

// solenoidal random-phase k^4 exp(-2(k/k0)^2) mode (same construction and counters as the oracle)
__device__ void k4_mode(int N, long long kx, long long ky, long long kz, double k0, unsigned long long seed, double2 out[3]) {
    const int Nzf = N / 2 + 1;
    const long long k2 = kx * kx + ky * ky + kz * kz;
    const long long cut = N / 3;
    out[0] = out[1] = out[2] = make_double2(0.0, 0.0);
    if (k2 == 0 || k2 > cut * cut) return;
    const long long i = kx >= 0 ? kx : kx + N, j = ky >= 0 ? ky : ky + N;
    const unsigned long long n = ((unsigned long long)i * N + (unsigned long long)j) * Nzf + (unsigned long long)kz;
    const uint32_t key0 = (uint32_t)seed, key1 = (uint32_t)(seed >> 32);
    uint32_t r0[4] = {(uint32_t)n, (uint32_t)(n >> 32), 0u, 0u};
    uint32_t r1[4] = {(uint32_t)n, (uint32_t)(n >> 32), 1u, 0u};
    philox4x32_10(r0, key0, key1);
    philox4x32_10(r1, key0, key1);
    const double th1 = 2.0 * kPi * u01(r0[0], r0[1]);
    const double th2 = 2.0 * kPi * u01(r0[2], r0[3]);
    const double phi = 2.0 * kPi * u01(r1[0], r1[1]);
    const double kk = sqrt((double)k2);
    const double Ek = (double)k2 * (double)k2 * exp(-2.0 * (kk / k0) * (kk / k0));
    const double A = sqrt(Ek / (4.0 * kPi * (double)k2));
    const double ar = A * cos(th1) * cos(phi), ai = A * sin(th1) * cos(phi);
    const double br = A * cos(th2) * sin(phi), bi = A * sin(th2) * sin(phi);
    double e1[3], e2[3];
    const long long kh2 = kx * kx + ky * ky;
    if (kh2 == 0) {
        e1[0] = 1; e1[1] = 0; e1[2] = 0; e2[0] = 0; e2[1] = 1; e2[2] = 0;
    } else {
        const double kh = sqrt((double)kh2);
        e1[0] = ky / kh; e1[1] = -kx / kh; e1[2] = 0.0;
        e2[0] = kx * kz / (kk * kh); e2[1] = ky * kz / (kk * kh); e2[2] = -(double)kh2 / (kk * kh);
    }
    for (int d = 0; d < 3; ++d) out[d] = make_double2(ar * e1[d] + br * e2[d], ai * e1[d] + bi * e2[d]);
}

// writes unscaled modes into the state; part[blockIdx] = block energy sum (half-spectrum weights)
__global__ void ic_k4_kernel(double2* __restrict__ uh, double* __restrict__ part, int N, int kmax, int NK, int NZP,
                             double k0, unsigned long long seed) {
    __shared__ double red[256];
    const long long plane = (long long)NK * NK * NZP;
    const long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    double e = 0.0;
    if (n < plane) {
        const int kz = (int)(n % NZP), ix = (int)((n / NZP) % NK), iy = (int)(n / ((long long)NZP * NK));
        const long long kx = k_of_cidx(ix, kmax, NK), ky = k_of_cidx(iy, kmax, NK);
        double2 v[3];
        if (kz > kmax) {
            v[0] = v[1] = v[2] = make_double2(0.0, 0.0);
        } else if (kz == 0 && (ky < 0 || (ky == 0 && kx < 0))) {
            k4_mode(N, -kx, -ky, 0, k0, seed, v);
            for (int d = 0; d < 3; ++d) v[d].y = -v[d].y;
        } else {
            k4_mode(N, kx, ky, kz, k0, seed, v);
        }
        const double wg = (kz == 0) ? 1.0 : 2.0;
        for (int d = 0; d < 3; ++d) {
            uh[d * plane + n] = v[d];
            e += wg * (v[d].x * v[d].x + v[d].y * v[d].y);
        }
    }
    const double s = block_sum(e, red);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// one block: thread t sums parts t, t + blockDim, ... in order, then a fixed tree (deterministic)
__global__ void sum_parts_kernel(const double* __restrict__ part, int n, double* __restrict__ out) {
    __shared__ double red[256];
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) s += part[i];
    s = block_sum(s, red);
    if (threadIdx.x == 0) out[0] = s;
}

__global__ void scale_state_kernel(double2* __restrict__ uh, long long n, const double* __restrict__ esum, double energy, double N3) {
    const double e_now = 0.5 * esum[0];
    const double s = (e_now > 0.0) ? sqrt(energy / e_now) * N3 : 0.0;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        uh[i].x *= s;
        uh[i].y *= s;
    }
}

}  // namespace nsg
