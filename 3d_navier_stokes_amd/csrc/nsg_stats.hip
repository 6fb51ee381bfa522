// nsg_stats.hip -- longitudinal velocity-increment statistics of real-space snapshots on the GPU
// (the reference's post-processing: PostProcessing/src/stats.c Precompute :31-150, ComputeStats
// :153-250, WriteStatsData :316-371, AllocateStatsObjects :404-423).
//
// For every separation r in {1,2,4,...,128,N/2} the increments u_x(i+r,j,k) - u_x(i,j,k),
// u_y(i,j+r,k) - u_y(i,j,k) and u_z(i,j,k+r) - u_z(i,j,k) with the shifted index inside the box
// (no periodic wrap) go into one accumulator per r: count, min, max and the central moments M2..M4
// (GSL's running statistics), then into a 1000-bin histogram on [min - 0.05, max + 0.05].
//
// One kernel per direction streams the AoS field once: a block stages a tile of T lines of the
// component along that direction in LDS (all N points of each line), then forms every increment
// of the tile for all nine r.  Blocks are persistent (a grid of a few blocks per CU walks all
// tiles), so the per-block partial moments and LDS histograms are flushed once.
//   moments: per-thread shifted power sums (shift = the block's first increment of that r), block
//            reduction, per-block partials merged with the pairwise central-moment update
//            (Chan et al. / Pebay) in a fixed order by incr_finalize_kernel -- deterministic;
//   histogram: the bin of x is the i with R(i) <= x < R(i+1), R(i) = ((n-i)/n) xmin + (i/n) xmax
//            evaluated without contraction exactly as the host computes the ranges (so every
//            count equals the sequential reference's), counted in LDS and flushed with 64-bit
//            atomics.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <string>
#include <vector>

#include "nsg_internal.hpp"

namespace {

constexpr int NINC = NSG_INCR_NUM;    // 9 separations
constexpr int NBIN = NSG_INCR_BINS;   // 1000 bins
constexpr int TB = 256;               // threads per block
constexpr int PSTRIDE = 8;            // doubles per partial: n, K, s1, s2, s3, s4, min, max

struct IncrArgs {
    const double* u;                 // AoS [N][N][N][3], or three planes [3][N][N][N] (soa)
    int soa;
    double scale;                    // multiplies every value read
    int N, T, lgT, ntiles;
    int inc[NINC];
    double* part;                    // moments: [gridDim.x][NINC][PSTRIDE] for this direction
    unsigned long long* counts;      // histogram: [NINC][NBIN]
    unsigned long long* outside;     // histogram: increments outside the ranges
    double lo[NINC], hi[NINC];       // histogram: range[0] and range[NBIN] per separation
    double inv[NINC];                // histogram: NBIN / (hi - lo), the linear map to bins
    double eps[NINC];                // histogram: its error bound vs the exact ranges (bin widths)
    const double* f;                 // histogram: f1[i] = (NBIN-i)/NBIN, f2[i] = i/NBIN (device)
};

// tile `tile` of direction DIR: LDS line t (0..T-1), point a (0..N-1).  Loads are issued in
// batches of LB per thread before their LDS stores, so LB HBM requests per thread are in flight.
// T is a power of two (lgT = log2 T).
template <int DIR, int NT>
__device__ __forceinline__ void load_tile(const IncrArgs& A, int tile, double* lds, int LD) {
    constexpr int LB = 8;
    const long long N = A.N;
    const int T = A.T, lgT = A.lgT;
    const int per = A.N >> lgT;
    const long long p0 = tile / per, p1 = (long long)(tile % per) << lgT;
    const int n = A.N << lgT;
    // DIR 2 walks idx = t N + a (a contiguous in memory): (t, a) advanced without dividing
    int t2 = 0, a2 = 0;
    if constexpr (DIR == 2) {
        t2 = threadIdx.x / A.N;
        a2 = threadIdx.x - t2 * A.N;
    }
    for (int base = 0; base < n; base += LB * NT) {
        double v[LB];
        int li[LB];
#pragma unroll
        for (int b = 0; b < LB; ++b) {
            const int idx = base + b * NT + threadIdx.x;
            long long off = 0;
            // point (i, j, k) of component DIR: AoS 3 ((i N + j) N + k) + DIR, SoA DIR N^3 + (i N + j) N + k
            long long pt;
            if constexpr (DIR == 2) {
                pt = (p0 * N + p1 + t2) * N + a2;
                li[b] = t2 * LD + a2;
                a2 += NT;
                while (a2 >= A.N) { a2 -= A.N; ++t2; }
            } else {
                const int t = idx & (T - 1), a = idx >> lgT;  // lines along i (DIR 0) or j (DIR 1)
                pt = DIR == 0 ? ((long long)a * N + p0) * N + p1 + t : (p0 * N + a) * N + p1 + t;
                li[b] = t * LD + a;
            }
            off = A.soa ? (long long)DIR * N * N * N + pt : pt * 3 + DIR;
            v[b] = 0.0;
            if (idx < n) v[b] = A.u[off];  // off is inside the field only for idx < n
        }
#pragma unroll
        for (int b = 0; b < LB; ++b)
            if (base + b * NT + (int)threadIdx.x < n) lds[li[b]] = A.scale * v[b];
    }
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}
__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    return v;
}

// ------------------------------------------------------------------------------------------
// moments pass: one wavefront per separation (NINC wavefronts per block), so a thread carries
// the 7 accumulators of one separation only and many wavefronts fit on a SIMD
// ------------------------------------------------------------------------------------------
constexpr int TBM = 64 * NINC;

template <int DIR>
__global__ void __launch_bounds__(TBM) incr_moments_kernel(IncrArgs A) {
    extern __shared__ double lds[];
    const int LD = A.N + 1;
    const int lane = threadIdx.x & 63, q = threadIdx.x >> 6;
    const int r = A.inc[q];
    const int len = A.N - r;  // increments per line (<= 0: none)
    double K = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0, s4 = 0.0, mn = INFINITY, mx = -INFINITY;
    bool first = true;
    long long tiles_here = 0;
    for (int tile = blockIdx.x; tile < A.ntiles; tile += gridDim.x) {
        __syncthreads();  // the previous tile is consumed
        load_tile<DIR, TBM>(A, tile, lds, LD);
        __syncthreads();
        ++tiles_here;
        if (len <= 0) continue;
        if (first) {
            K = lds[r] - lds[0];  // shift: this block's first increment of the separation
            first = false;
        }
        // pair p = t len + a walked with stride 64 ((t, a) advanced without dividing)
        int t = lane / len, a = lane - t * len;
        while (t < A.T) {
            const int base = t * LD + a;
            const double d = lds[base + r] - lds[base];
            const double y = d - K, y2 = y * y;
            s1 += y;
            s2 += y2;
            s3 += y2 * y;
            s4 += y2 * y2;
            mn = fmin(mn, d);
            mx = fmax(mx, d);
            a += 64;
            while (a >= len) { a -= len; ++t; }
        }
    }
    const double v1 = wave_sum(s1), v2 = wave_sum(s2), v3 = wave_sum(s3), v4 = wave_sum(s4);
    const double vmn = wave_min(mn), vmx = wave_max(mx);
    if (lane == 0) {
        const double cnt = len > 0 ? (double)len * A.T * (double)tiles_here : 0.0;
        double* dst = A.part + ((long long)blockIdx.x * NINC + q) * PSTRIDE;
        dst[0] = cnt;
        dst[1] = K;
        dst[2] = v1;
        dst[3] = v2;
        dst[4] = v3;
        dst[5] = v4;
        dst[6] = cnt > 0.0 ? vmn : 0.0;
        dst[7] = cnt > 0.0 ? vmx : 0.0;
    }
}

// running statistics of one set: n, mean, M2, M3, M4, min, max
struct Mom {
    double n, mean, M2, M3, M4, mn, mx;
};

__device__ __forceinline__ Mom from_part(const double* p) {
    Mom m;
    m.n = p[0];
    if (m.n <= 0.0) return Mom{0, 0, 0, 0, 0, 0, 0};
    const double a = p[2] / m.n;  // mean of the shifted values
    m.mean = p[1] + a;
    m.M2 = p[3] - m.n * a * a;
    m.M3 = p[4] - 3.0 * a * p[3] + 2.0 * m.n * a * a * a;
    m.M4 = p[5] - 4.0 * a * p[4] + 6.0 * a * a * p[3] - 3.0 * m.n * a * a * a * a;
    m.mn = p[6];
    m.mx = p[7];
    return m;
}

// pairwise update of the central moments of A and B (Chan et al. 1979; Pebay 2008 for M3, M4)
__device__ __forceinline__ Mom merge(const Mom& A, const Mom& B) {
    if (B.n <= 0.0) return A;
    if (A.n <= 0.0) return B;
    Mom m;
    const double na = A.n, nb = B.n, n = na + nb, d = B.mean - A.mean, d2 = d * d;
    m.n = n;
    m.mean = A.mean + d * nb / n;
    m.M2 = A.M2 + B.M2 + d2 * na * nb / n;
    m.M3 = A.M3 + B.M3 + d2 * d * na * nb * (na - nb) / (n * n) + 3.0 * d * (na * B.M2 - nb * A.M2) / n;
    m.M4 = A.M4 + B.M4 + d2 * d2 * na * nb * (na * na - na * nb + nb * nb) / (n * n * n) +
           6.0 * d2 * (na * na * B.M2 + nb * nb * A.M2) / (n * n) + 4.0 * d * (na * B.M3 - nb * A.M3) / n;
    m.mn = fmin(A.mn, B.mn);
    m.mx = fmax(A.mx, B.mx);
    return m;
}

// acc[q] <- acc[q] merged with every partial of separation q (three directions' partials); one
// block per separation, fixed merge order
__global__ void __launch_bounds__(TB) incr_finalize_kernel(double* acc, const double* part, int nparts) {
    __shared__ Mom sh[TB];
    const int q = blockIdx.x;
    Mom m{0, 0, 0, 0, 0, 0, 0};
    for (int j = threadIdx.x; j < nparts; j += TB) m = merge(m, from_part(part + ((long long)j * NINC + q) * PSTRIDE));
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int s = TB / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] = merge(sh[threadIdx.x], sh[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double* a = acc + q * 8;
        const Mom old{a[0], a[3], a[4], a[5], a[6], a[1], a[2]};
        const Mom r = merge(old, sh[0]);
        a[0] = r.n; a[1] = r.mn; a[2] = r.mx; a[3] = r.mean; a[4] = r.M2; a[5] = r.M3; a[6] = r.M4;
    }
}

// ------------------------------------------------------------------------------------------
// histogram pass
// ------------------------------------------------------------------------------------------
// range i of the uniform histogram on [lo, hi]: f1[i] lo + f2[i] hi with f1[i] = (n-i)/n and
// f2[i] = i/n (tables made on the host), evaluated without contraction as the host does
__device__ __forceinline__ double range_at(const double* f, int i, double lo, double hi) {
#pragma clang fp contract(off)
    return f[i] * lo + f[NBIN + 1 + i] * hi;
}

// one wavefront per separation, as the moments pass
template <int DIR>
__global__ void __launch_bounds__(TBM) incr_hist_kernel(IncrArgs A) {
    extern __shared__ double lds[];
    const int LD = A.N + 1;
    const double* f = A.f;                                     // f1, f2 (global: rarely read)
    unsigned int* hist = (unsigned int*)(lds + (size_t)A.T * LD);  // [NINC][NBIN]
    for (int b = threadIdx.x; b < NINC * NBIN; b += TBM) hist[b] = 0u;
    const int lane = threadIdx.x & 63, q = threadIdx.x >> 6;
    const int r = A.inc[q], len = A.N - r;
    const double lo = A.lo[q], hi = A.hi[q], inv = A.inv[q], eps = A.eps[q];
    unsigned int* hq = hist + q * NBIN;
    unsigned long long out = 0;
    for (int tile = blockIdx.x; tile < A.ntiles; tile += gridDim.x) {
        __syncthreads();
        load_tile<DIR, TBM>(A, tile, lds, LD);
        __syncthreads();
        if (len <= 0) continue;
        int t = lane / len, a = lane - t * len;
        while (t < A.T) {
            const int base = t * LD + a;
            const double x = lds[base + r] - lds[base];
            a += 64;
            while (a >= len) { a -= len; ++t; }
            if (!(x >= lo && x < hi)) {  // outside [range[0], range[n]) or NaN
                ++out;
                continue;
            }
            // the linear map decides unless x lies within the rounding margin of a bin edge
            // (margin = a bound on the error of the map vs the exact ranges, in bin widths); there
            // the exact ranges decide
            const double u = (x - lo) * inv;
            int i = (int)u;
            const double fr = u - (double)i;
            if (i > NBIN - 1 || fr < eps || fr > 1.0 - eps) {
                i = i > NBIN - 1 ? NBIN - 1 : i;
                while (i > 0 && x < range_at(f, i, lo, hi)) --i;
                while (i < NBIN - 1 && x >= range_at(f, i + 1, lo, hi)) ++i;
            }
            atomicAdd(&hq[i], 1u);
        }
    }
    __syncthreads();
    for (int b = threadIdx.x; b < NINC * NBIN; b += TBM)
        if (hist[b]) atomicAdd(&A.counts[b], (unsigned long long)hist[b]);
    double o = wave_sum((double)out);
    if ((threadIdx.x & 63) == 0 && o > 0.0) atomicAdd(A.outside, (unsigned long long)o);
}

}  // namespace

// ------------------------------------------------------------------------------------------
// object and C ABI
// ------------------------------------------------------------------------------------------
struct nsg_incr {
    int N = 0, device = 0, T = 1, Th = 1, ncu = 256;  // lines per tile: moments, histogram
    double* ftab = nullptr;                           // f1, f2 of the uniform ranges
    int inc[NINC] = {};
    hipStream_t s = nullptr;
    double* u = nullptr;                   // device copy of a host snapshot
    double* part = nullptr;                // partial moments of the three directions
    int grid = 1;                          // persistent blocks per direction
    double* acc = nullptr;                 // [NINC][8]: n, min, max, mean, M2, M3, M4, -
    unsigned long long* counts = nullptr;  // [NINC][NBIN], then the outside counter
    double ranges[NINC][NBIN + 1] = {};
    bool have_ranges = false;
    std::string err;
};

static int ifail(nsg_incr* s, int code, const char* msg) {
    if (s) s->err = msg;
    return code;
}
#define IHIP(s, expr)                                                                                \
    do {                                                                                             \
        hipError_t e_ = (expr);                                                                      \
        if (e_ != hipSuccess) return ifail((s), NSG_ERR_DEVICE, (std::string(#expr) + ": " + hipGetErrorString(e_)).c_str()); \
    } while (0)

// lines per tile: a divisor of N, with the tile (T (N+1) doubles) within `kb` KB
static int tile_lines(int N, int kb) {
    for (int T : {16, 8, 4, 2, 1})
        if (N % T == 0 && (size_t)T * (N + 1) * sizeof(double) <= (size_t)kb * 1024) return T;
    return 1;
}
static int ilog2(int T) {
    int l = 0;
    while ((1 << l) < T) ++l;
    return l;
}
// histogram kernel LDS: tile and NINC x NBIN counters (three blocks per CU fit)
static size_t hist_lds_bytes(int N, int Th) {
    return (size_t)Th * (N + 1) * sizeof(double) + (size_t)NINC * NBIN * sizeof(unsigned int);
}

static IncrArgs base_args(const nsg_incr* s, const double* u, double scale, int soa) {
    IncrArgs a{};
    a.u = u;
    a.soa = soa;
    a.scale = scale;
    a.N = s->N;
    a.T = s->T;
    a.lgT = ilog2(s->T);
    a.ntiles = s->N * (s->N / s->T);
    for (int q = 0; q < NINC; ++q) a.inc[q] = s->inc[q];
    return a;
}

// moments or histogram pass over a device field on `st`
static int run_pass(nsg_incr* s, const double* u, double scale, int hist, hipStream_t st, int soa = 0) {
    IncrArgs a = base_args(s, u, scale, soa);
    const size_t tile_bytes = (size_t)s->T * (s->N + 1) * sizeof(double);
    if (!hist) {
        for (int dir = 0; dir < 3; ++dir) {
            a.part = s->part + (size_t)dir * s->grid * NINC * PSTRIDE;
            if (dir == 0) hipLaunchKernelGGL(incr_moments_kernel<0>, dim3(s->grid), dim3(TBM), tile_bytes, st, a);
            if (dir == 1) hipLaunchKernelGGL(incr_moments_kernel<1>, dim3(s->grid), dim3(TBM), tile_bytes, st, a);
            if (dir == 2) hipLaunchKernelGGL(incr_moments_kernel<2>, dim3(s->grid), dim3(TBM), tile_bytes, st, a);
            IHIP(s, hipGetLastError());
        }
        hipLaunchKernelGGL(incr_finalize_kernel, dim3(NINC), dim3(TB), 0, st, s->acc, (const double*)s->part, 3 * s->grid);
        IHIP(s, hipGetLastError());
    } else {
        if (!s->have_ranges) return ifail(s, NSG_ERR_INVALID, "nsg_incr_set_ranges must follow the precompute passes");
        a.counts = s->counts;
        a.outside = s->counts + NINC * NBIN;
        for (int q = 0; q < NINC; ++q) {
            a.lo[q] = s->ranges[q][0];
            a.hi[q] = s->ranges[q][NBIN];
            a.inv[q] = NBIN / (a.hi[q] - a.lo[q]);
            // each range R(i) = f1 lo + f2 hi carries at most 2 eps (|lo| + |hi|) of rounding, the
            // map (x - lo) inv at most 2 eps NBIN bin widths; four times their sum is the margin
            a.eps[q] = 16.0 * DBL_EPSILON * NBIN * (1.0 + (std::fabs(a.lo[q]) + std::fabs(a.hi[q])) / (a.hi[q] - a.lo[q]));
        }
        a.f = s->ftab;
        a.T = s->Th;
        a.lgT = ilog2(s->Th);
        a.ntiles = s->N * (s->N / s->Th);
        const size_t bytes = hist_lds_bytes(s->N, s->Th);
        // resident blocks per CU: as many as the 160 KB LDS holds (two with 8-line tiles at 512^3)
        const int per_cu = std::max(1, (int)((160 * 1024) / hist_lds_bytes(s->N, s->Th)));
        const int hg = std::max(1, std::min(a.ntiles, std::min(3, per_cu) * s->ncu));
        hipLaunchKernelGGL(incr_hist_kernel<0>, dim3(hg), dim3(TBM), bytes, st, a);
        hipLaunchKernelGGL(incr_hist_kernel<1>, dim3(hg), dim3(TBM), bytes, st, a);
        hipLaunchKernelGGL(incr_hist_kernel<2>, dim3(hg), dim3(TBM), bytes, st, a);
        IHIP(s, hipGetLastError());
    }
    IHIP(s, hipStreamSynchronize(st));
    if (hist) {
        unsigned long long outside = 0;
        IHIP(s, hipMemcpy(&outside, s->counts + NINC * NBIN, sizeof outside, hipMemcpyDeviceToHost));
        if (outside) {
            IHIP(s, hipMemsetAsync(s->counts + NINC * NBIN, 0, sizeof outside, st));
            IHIP(s, hipStreamSynchronize(st));
            return ifail(s, NSG_ERR_INVALID,
                         (std::to_string(outside) + " increments fall outside the bin ranges (the reference exits here)").c_str());
        }
    }
    return NSG_OK;
}

int nsg_incr_create(int N, int device, nsg_incr** out) {
    if (!out) return NSG_ERR_INVALID;
    *out = nullptr;
    nsg_incr* s = new (std::nothrow) nsg_incr();
    if (!s) return NSG_ERR_OOM;
    *out = s;
    if (N < 4 || N > 1024) return ifail(s, NSG_ERR_INVALID, "N must be in [4, 1024]");
    s->N = N;
    s->device = device;
    s->T = tile_lines(N, 33);
    s->Th = tile_lines(N, 33);
    const int inc[NINC] = {1, 2, 4, 8, 16, 32, 64, 128, N / 2};
    std::copy(inc, inc + NINC, s->inc);
    IHIP(s, hipSetDevice(device));
    hipDeviceProp_t prop;
    IHIP(s, hipGetDeviceProperties(&prop, device));
    s->ncu = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    const int ntiles = N * (N / s->T);
    s->grid = std::max(1, std::min(ntiles, 3 * s->ncu));  // three resident moment blocks per CU (27 waves)
    const int hist_lds = (int)hist_lds_bytes(N, s->Th);
    IHIP(s, hipFuncSetAttribute((const void*)&incr_hist_kernel<0>, hipFuncAttributeMaxDynamicSharedMemorySize, hist_lds));
    IHIP(s, hipFuncSetAttribute((const void*)&incr_hist_kernel<1>, hipFuncAttributeMaxDynamicSharedMemorySize, hist_lds));
    IHIP(s, hipFuncSetAttribute((const void*)&incr_hist_kernel<2>, hipFuncAttributeMaxDynamicSharedMemorySize, hist_lds));
    IHIP(s, hipStreamCreateWithFlags(&s->s, hipStreamNonBlocking));
    IHIP(s, hipMalloc(&s->part, sizeof(double) * 3 * s->grid * NINC * PSTRIDE));
    IHIP(s, hipMalloc(&s->acc, sizeof(double) * NINC * 8));
    IHIP(s, hipMalloc(&s->counts, sizeof(unsigned long long) * (NINC * NBIN + 1)));
    {
        // gsl_histogram_set_ranges_uniform's weights, the same doubles the host ranges use
        std::vector<double> f(2 * (NBIN + 1));
        for (int b = 0; b <= NBIN; ++b) {
            f[b] = (double)(NBIN - b) / (double)NBIN;
            f[NBIN + 1 + b] = (double)b / (double)NBIN;
        }
        IHIP(s, hipMalloc(&s->ftab, sizeof(double) * f.size()));
        IHIP(s, hipMemcpy(s->ftab, f.data(), sizeof(double) * f.size(), hipMemcpyHostToDevice));
    }
    IHIP(s, hipMemset(s->acc, 0, sizeof(double) * NINC * 8));
    IHIP(s, hipMemset(s->counts, 0, sizeof(unsigned long long) * (NINC * NBIN + 1)));
    IHIP(s, hipDeviceSynchronize());
    return NSG_OK;
}

void nsg_incr_destroy(nsg_incr* s) {
    if (!s) return;
    (void)hipSetDevice(s->device);
    for (void* p : {(void*)s->u, (void*)s->part, (void*)s->acc, (void*)s->counts, (void*)s->ftab})
        if (p) (void)hipFree(p);
    if (s->s) (void)hipStreamDestroy(s->s);
    delete s;
}

const char* nsg_incr_last_error(const nsg_incr* s) { return s ? s->err.c_str() : "null handle"; }

int nsg_incr_increments(const nsg_incr* s, int* r) {
    if (!s || !r) return NSG_ERR_INVALID;
    std::copy(s->inc, s->inc + NINC, r);
    return NSG_OK;
}

static int host_pass(nsg_incr* s, const double* u, int hist) {
    if (!s || !u) return NSG_ERR_INVALID;
    IHIP(s, hipSetDevice(s->device));
    const size_t bytes = sizeof(double) * 3 * (size_t)s->N * s->N * s->N;
    if (!s->u) IHIP(s, hipMalloc(&s->u, bytes));
    IHIP(s, hipMemcpyAsync(s->u, u, bytes, hipMemcpyHostToDevice, s->s));
    return run_pass(s, s->u, 1.0, hist, s->s);
}

static int state_pass(nsg_incr* s, nsg_ctx* c, double scale, int hist) {
    if (!s || !c) return NSG_ERR_INVALID;
    if (c->nbox != 1 || c->N != s->N || c->p.device != s->device)
        return ifail(s, NSG_ERR_INVALID, "the context must hold one box of the same N on the same device");
    IHIP(s, hipSetDevice(s->device));
    // the context's c2r writes the field as three planes: each direction's kernel then streams
    // only its own component
    const double* field = nullptr;
    if (nsgi_realspace(c, &field, 1) != NSG_OK) return ifail(s, NSG_ERR_DEVICE, c->err.c_str());
    return run_pass(s, field, scale, hist, c->s, 1);
}

int nsg_incr_precompute(nsg_incr* s, const double* u) { return host_pass(s, u, 0); }
int nsg_incr_histogram(nsg_incr* s, const double* u) { return host_pass(s, u, 1); }
int nsg_incr_precompute_state(nsg_incr* s, nsg_ctx* c, double scale) { return state_pass(s, c, scale, 0); }
int nsg_incr_histogram_state(nsg_incr* s, nsg_ctx* c, double scale) { return state_pass(s, c, scale, 1); }

static int download_acc(nsg_incr* s, double* acc) {
    IHIP(s, hipSetDevice(s->device));
    IHIP(s, hipMemcpy(acc, s->acc, sizeof(double) * NINC * 8, hipMemcpyDeviceToHost));
    return NSG_OK;
}

int nsg_incr_set_ranges(nsg_incr* s) {
    if (!s) return NSG_ERR_INVALID;
    double acc[NINC * 8];
    int rc = download_acc(s, acc);
    if (rc) return rc;
    for (int q = 0; q < NINC; ++q) {
        const double xmin = acc[q * 8 + 1] - 0.05, xmax = acc[q * 8 + 2] + 0.05;
        if (!std::isfinite(xmin) || !std::isfinite(xmax) || !std::isfinite(acc[q * 8 + 3]))
            return ifail(s, NSG_ERR_NONFINITE, "non-finite increments in the precompute passes");
        // gsl_histogram_set_ranges_uniform
        for (int b = 0; b <= NBIN; ++b) {
            const double f1 = (double)(NBIN - b) / (double)NBIN;
            const double f2 = (double)b / (double)NBIN;
            s->ranges[q][b] = f1 * xmin + f2 * xmax;
        }
    }
    // the passes run on non-blocking streams: clear on the object's stream and wait
    IHIP(s, hipMemsetAsync(s->counts, 0, sizeof(unsigned long long) * (NINC * NBIN + 1), s->s));
    IHIP(s, hipStreamSynchronize(s->s));
    s->have_ranges = true;
    return NSG_OK;
}

int nsg_incr_results(nsg_incr* s, double* stats, double* ranges, double* counts) {
    if (!s) return NSG_ERR_INVALID;
    double acc[NINC * 8];
    int rc = download_acc(s, acc);
    if (rc) return rc;
    if (stats) {
        for (int q = 0; q < NINC; ++q) {
            const double* a = acc + q * 8;
            const double n = a[0];
            double* o = stats + 6 * q;
            // gsl_rstat_min/max/mean/sd/skew/kurtosis
            o[0] = n > 0.0 ? a[1] : 0.0;
            o[1] = n > 0.0 ? a[2] : 0.0;
            o[2] = a[3];
            o[3] = n > 1.0 ? std::sqrt(a[4] / (n - 1.0)) : 0.0;
            o[4] = n > 0.0 ? (std::pow(n - 1.0, 1.5) / n) * a[5] / std::pow(a[4], 1.5) : 0.0;
            o[5] = n > 0.0 ? ((n - 1.0) / n) * (n - 1.0) * a[6] / (a[4] * a[4]) - 3.0 : 0.0;
        }
    }
    if (ranges) {
        if (!s->have_ranges) return ifail(s, NSG_ERR_INVALID, "no bin ranges yet (nsg_incr_set_ranges)");
        std::copy(&s->ranges[0][0], &s->ranges[0][0] + NINC * (NBIN + 1), ranges);
    }
    if (counts) {
        std::vector<unsigned long long> h(NINC * NBIN);
        IHIP(s, hipMemcpy(h.data(), s->counts, sizeof(unsigned long long) * NINC * NBIN, hipMemcpyDeviceToHost));
        for (int b = 0; b < NINC * NBIN; ++b) counts[b] = (double)h[b];
    }
    return NSG_OK;
}

int nsg_incr_reset(nsg_incr* s) {
    if (!s) return NSG_ERR_INVALID;
    IHIP(s, hipSetDevice(s->device));
    IHIP(s, hipMemsetAsync(s->acc, 0, sizeof(double) * NINC * 8, s->s));
    IHIP(s, hipMemsetAsync(s->counts, 0, sizeof(unsigned long long) * (NINC * NBIN + 1), s->s));
    IHIP(s, hipStreamSynchronize(s->s));
    s->have_ranges = false;
    return NSG_OK;
}

// times `iters` passes over the last host snapshot (they accumulate: reset afterwards)
int nsg_incr_time_pass(nsg_incr* s, int hist, int iters, double* ms) {
    if (!s || iters <= 0 || !ms || !s->u) return NSG_ERR_INVALID;
    IHIP(s, hipSetDevice(s->device));
    hipEvent_t e0, e1;
    IHIP(s, hipEventCreate(&e0));
    IHIP(s, hipEventCreate(&e1));
    int rc = run_pass(s, s->u, 1.0, hist, s->s);  // untimed first launch
    if (rc) return rc;
    IHIP(s, hipEventRecord(e0, s->s));
    for (int i = 0; i < iters && rc == NSG_OK; ++i) rc = run_pass(s, s->u, 1.0, hist, s->s);
    IHIP(s, hipEventRecord(e1, s->s));
    IHIP(s, hipEventSynchronize(e1));
    float t = 0.f;
    IHIP(s, hipEventElapsedTime(&t, e0, e1));
    *ms = t / iters;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}
