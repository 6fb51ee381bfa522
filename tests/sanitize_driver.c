/* CPU sanitizer driver (test infrastructure): runs every oracle entry point on small grids so an
 * AddressSanitizer / UndefinedBehaviorSanitizer build of oracle/nsg_oracle.c and
 * oracle/stats_oracle.c can report out-of-bounds accesses, leaks and undefined arithmetic.  Built
 * and run by tests/test_sanitizers.py; prints "ok" and exits 0 when every check holds. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "nsg_oracle.h"

static int check(int cond, const char* what, int N) {
    if (!cond) fprintf(stderr, "FAILED: %s (N=%d)\n", what, N);
    return cond ? 0 : 1;
}

static int run(int N) {
    const int nzf = N / 2 + 1, ns = oracle_n_spect(N);
    const size_t nspec = (size_t)N * N * nzf * 3 * 2, nreal = (size_t)N * N * N * 3;
    double* uh = calloc(nspec, sizeof(double));
    double* u2 = calloc(nspec, sizeof(double));
    double* nl = calloc(nspec, sizeof(double));
    double* ur = calloc(nreal, sizeof(double));
    double* sc = calloc((size_t)N * N * N, sizeof(double));
    double* cc = calloc((size_t)N * N * nzf * 2, sizeof(double));
    double* ek = calloc((size_t)ns, sizeof(double));
    double* ens = calloc((size_t)ns, sizeof(double));
    double tot[5], acc[ORACLE_NUM_INCR * 7], rng[ORACLE_NUM_INCR * (ORACLE_N_BINS + 1)];
    double cnt[ORACLE_NUM_INCR * ORACLE_N_BINS], st[ORACLE_NUM_INCR * 6];
    int kx[4096], ky[4096], kz[4096], incr[ORACLE_NUM_INCR], bad = 0;

    oracle_wavenumbers(N, kx, ky, kz);
    for (size_t i = 0; i < (size_t)N * N * N; ++i) sc[i] = sin(0.37 * (double)i) + 0.1 * (double)(i % 7);
    oracle_r2c3d(N, sc, cc);
    double* back = calloc((size_t)N * N * N, sizeof(double));
    oracle_c2r3d(N, cc, back);
    double err = 0.0;
    for (size_t i = 0; i < (size_t)N * N * N; ++i) err = fmax(err, fabs(back[i] / ((double)N * N * N) - sc[i]));
    bad += check(err < 1e-12, "c2r(r2c(x)) = N^3 x", N);
    double line[2 * 64];
    for (int i = 0; i < 2 * N; ++i) line[i] = cos(0.3 * i);
    oracle_fft1d(N, line, -1);
    oracle_fft1d(N, line, +1);
    bad += check(fabs(line[2] / N - cos(0.6)) < 1e-12, "fft1d round trip", N);

    oracle_ic_taylor_green(N, uh);
    oracle_measures(N, uh, 0.01, 1, 0, tot, ek, ens);
    bad += check(fabs(tot[0] - M_PI * M_PI * M_PI) < 1e-10, "Taylor-Green energy pi^3", N);
    oracle_measures(N, uh, 0.01, 1, 1, tot, ek, ens);
    oracle_ic_shapiro(N, 0.01, 0.0, uh);
    oracle_ic_random_ref(N, uh);
    oracle_ic_k4(N, 2.0, 0.5, 7, uh);
    oracle_dealias(N, uh);
    oracle_curl(N, uh, u2);
    oracle_rhs(N, uh, nl);
    oracle_forcing(N, uh, 0.1, 2.5, nl);
    memcpy(u2, uh, nspec * sizeof(double));
    for (int sch = 0; sch < 4; ++sch) oracle_step(N, u2, 0.01, 1 + (sch & 1), 1e-3, sch, sch == 1 ? 0.1 : 0.0, 2.5);
    double e = oracle_step_rk5(N, uh, u2, 0.01, 1, 1e-3, 7, 0, 0.0, 2.5);
    bad += check(e >= 0.0 && isfinite(e), "embedded error norm finite", N);
    (void)oracle_step_rk5(N, uh, u2, 0.01, 2, 1e-3, 6, 1, 0.1, 2.5);
    double dt = 1e-3, t = 0.0;
    long rej = 0;
    double last = 0.0;
    bad += check(oracle_dp5_solve(N, u2, 0.01, 1, &dt, &t, 2, 0.0, 2.5, &rej, &last) == 0, "dp5 two steps", N);
    bad += check(oracle_dp5_factor(1.0) == 0.8, "dp5 factor at err 1", N);
    oracle_realspace(N, uh, 0, ur);
    oracle_realspace(N, uh, 1, ur);
    bad += check(oracle_max_speed_sum(N, uh) > 0.0 && oracle_cfl_dt(N, uh, 0.5) > 0.0, "max speed / CFL dt", N);
    uint32_t c4[4] = {1, 2, 3, 4}, k2[2] = {5, 6}, o4[4];
    oracle_philox4x32_10(c4, k2, o4);

    oracle_realspace(N, uh, 0, ur);
    memset(acc, 0, sizeof acc);
    memset(cnt, 0, sizeof cnt);
    oracle_incr_list(N, incr);
    oracle_incr_moments(N, ur, acc);
    oracle_incr_ranges(acc, rng);
    bad += check(oracle_incr_histogram(N, ur, rng, cnt) == 0, "every increment inside the ranges", N);
    oracle_incr_stats(acc, st);

    free(uh); free(u2); free(nl); free(ur); free(sc); free(cc); free(ek); free(ens); free(back);
    return bad;
}

int main(void) {
    int bad = 0;
    oracle_set_threads(1);
    const int sizes[] = {8, 12, 16};
    for (int i = 0; i < 3; ++i) bad += run(sizes[i]);
    if (bad) return 1;
    printf("ok\n");
    return 0;
}
