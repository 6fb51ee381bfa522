/*
 * stats_oracle.c -- CPU restatement of the reference post-processing increment statistics
 * (TEST INFRASTRUCTURE; loaded only by tests/ and bench.py's cpu_baseline leg).
 *
 * Reference: PostProcessing/src/stats.c -- Precompute (:31-150), ComputeStats (:153-250), the
 * statistics and histograms written by WriteStatsData (:316-371), AllocateStatsObjects
 * (:404-423); constants PostProcessing/src/data_types.h:49-51 (N_BINS 1000, NUM_INCR 9,
 * INCR_TYPES 1).
 *
 * The reference accumulates with GSL (PostProcessing/Makefile:19 links the system -lgsl; the
 * version is not pinned and GSL is absent here).  Its published algorithms are restated:
 *   gsl_rstat_add: n += 1; delta = x - mean; delta_n = delta / n; term1 = delta delta_n (n - 1);
 *     mean += delta_n;
 *     M4 += term1 delta_n^2 (n^2 - 3n + 3) + 6 delta_n^2 M2 - 4 delta_n M3;
 *     M3 += term1 delta_n (n - 2) - 3 delta_n M2;  M2 += term1;  min/max (first value sets both).
 *   gsl_rstat_sd = sqrt(M2 / (n - 1)) (0 for n <= 1); skew = ((n-1)^1.5 / n) M3 / M2^1.5;
 *     kurtosis = ((n-1)^2 / n) M4 / M2^2 - 3 (0 for n == 0); an empty workspace reports 0 for
 *     min and max (gsl_rstat_reset).
 *   gsl_histogram_set_ranges_uniform: range[i] = ((n-i)/n) xmin + (i/n) xmax, i = 0..n.
 *   gsl_histogram_increment: the bin i with range[i] <= x < range[i+1]; x outside
 *     [range[0], range[n]) is an error (the reference exits, :191-195).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "nsg_oracle.h"

#define NI ORACLE_NUM_INCR
#define NB ORACLE_N_BINS

/* AllocateStatsObjects, PostProcessing/src/stats.c:407-413 */
void oracle_incr_list(int N, int* r) {
    const int inc[NI] = {1, 2, 4, 8, 16, 32, 64, 128, N / 2};
    memcpy(r, inc, sizeof inc);
}

/* acc[7] = {n, min, max, mean, M2, M3, M4} */
static void rstat_add(double x, double* a) {
    const double delta = x - a[3];
    if (a[0] == 0.0) {
        a[1] = x;
        a[2] = x;
    } else {
        if (x < a[1]) a[1] = x;
        if (x > a[2]) a[2] = x;
    }
    const double n = (a[0] += 1.0);
    const double delta_n = delta / n;
    const double delta_nsq = delta_n * delta_n;
    const double term1 = delta * delta_n * (n - 1.0);
    a[3] += delta_n;
    a[6] += term1 * delta_nsq * (n * n - 3.0 * n + 3.0) + 6.0 * delta_nsq * a[4] - 4.0 * delta_n * a[5];
    a[5] += term1 * delta_n * (n - 2.0) - 3.0 * delta_n * a[4];
    a[4] += term1;
}

/* Precompute's loop over one snapshot u [N][N][N][3] (PostProcessing/src/stats.c:62-110): for every
   point in (i, j, k) order and every increment r, the longitudinal increments in x, y and z when
   the shifted index stays inside the box (no periodic wrap) */
void oracle_incr_moments(int N, const double* u, double* acc /* [NI][7] */) {
    int inc[NI];
    oracle_incr_list(N, inc);
    const long Nl = N;
    for (long i = 0; i < Nl; ++i)
        for (long j = 0; j < Nl; ++j)
            for (long k = 0; k < Nl; ++k) {
                const long p = Nl * (i * Nl + j) + k;
                for (int q = 0; q < NI; ++q) {
                    const long r = inc[q];
                    if (i + r < Nl) rstat_add(u[3 * (Nl * ((i + r) * Nl + j) + k) + 0] - u[3 * p + 0], acc + 7 * q);
                    if (j + r < Nl) rstat_add(u[3 * (Nl * (i * Nl + j + r) + k) + 1] - u[3 * p + 1], acc + 7 * q);
                    if (k + r < Nl) rstat_add(u[3 * (Nl * (i * Nl + j) + k + r) + 2] - u[3 * p + 2], acc + 7 * q);
                }
            }
}

/* PostProcessing/src/stats.c:118-132 with gsl_histogram_set_ranges_uniform */
void oracle_incr_ranges(const double* acc, double* ranges /* [NI][NB+1] */) {
    for (int q = 0; q < NI; ++q) {
        const double xmin = acc[7 * q + 1] - 0.05, xmax = acc[7 * q + 2] + 0.05;
        for (int b = 0; b <= NB; ++b) {
            const double f1 = (double)(NB - b) / (double)NB;
            const double f2 = (double)b / (double)NB;
            ranges[q * (NB + 1) + b] = f1 * xmin + f2 * xmax;
        }
    }
}

/* gsl_histogram's bin search (linear guess, then bisection); -1 outside [range[0], range[n]) */
static long find_bin(const double* range, double x) {
    if (!(x >= range[0]) || x >= range[NB]) return -1;
    const double u = (x - range[0]) / (range[NB] - range[0]);
    long i = (long)(u * NB);
    if (i >= 0 && i < NB && x >= range[i] && x < range[i + 1]) return i;
    long lo = 0, hi = NB;
    while (hi - lo > 1) {
        const long mid = (hi + lo) / 2;
        if (x >= range[mid]) lo = mid;
        else hi = mid;
    }
    return lo;
}

/* ComputeStats (PostProcessing/src/stats.c:175-246): counts [NI][NB] += this snapshot's
   histogram; returns the number of increments outside the ranges (each one is a reference exit) */
long oracle_incr_histogram(int N, const double* u, const double* ranges, double* counts) {
    int inc[NI];
    oracle_incr_list(N, inc);
    const long Nl = N;
    long outside = 0;
    for (long i = 0; i < Nl; ++i)
        for (long j = 0; j < Nl; ++j)
            for (long k = 0; k < Nl; ++k) {
                const long p = Nl * (i * Nl + j) + k;
                for (int q = 0; q < NI; ++q) {
                    const long r = inc[q];
                    const double* rg = ranges + q * (NB + 1);
                    double v[3];
                    int nv = 0;
                    if (i + r < Nl) v[nv++] = u[3 * (Nl * ((i + r) * Nl + j) + k) + 0] - u[3 * p + 0];
                    if (j + r < Nl) v[nv++] = u[3 * (Nl * (i * Nl + j + r) + k) + 1] - u[3 * p + 1];
                    if (k + r < Nl) v[nv++] = u[3 * (Nl * (i * Nl + j) + k + r) + 2] - u[3 * p + 2];
                    for (int m = 0; m < nv; ++m) {
                        const long b = find_bin(rg, v[m]);
                        if (b < 0) ++outside;
                        else counts[q * NB + b] += 1.0;
                    }
                }
            }
    return outside;
}

/* the VelocityIncrementStats dataset (PostProcessing/src/stats.c:322-331): [NI][6] = min, max,
   mean, sd, skew, kurtosis with GSL's definitions */
void oracle_incr_stats(const double* acc, double* out /* [NI][6] */) {
    for (int q = 0; q < NI; ++q) {
        const double* a = acc + 7 * q;
        const double n = a[0];
        double* o = out + 6 * q;
        o[0] = a[1];
        o[1] = a[2];
        o[2] = a[3];
        o[3] = n > 1.0 ? sqrt(a[4] / (n - 1.0)) : 0.0;
        o[4] = n > 0.0 ? (pow(n - 1.0, 1.5) / n) * a[5] / pow(a[4], 1.5) : 0.0;
        o[5] = n > 0.0 ? ((n - 1.0) / n) * (n - 1.0) * a[6] / (a[4] * a[4]) - 3.0 : 0.0;
    }
}
