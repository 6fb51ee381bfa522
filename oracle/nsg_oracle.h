/*
 * nsg_oracle.h -- CPU FP64 restatement of the reference pseudo-spectral time-step path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the GPU library: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The product path
 * (libnsg.so) never links, loads or calls it.
 *
 * Conventions (all follow Solver/src of the reference):
 *   - cubic grid N^3 on [0,2pi)^3, x_i = 2 pi i / N            (Solver/src/solver.c:1777-1823)
 *   - spectral arrays are interleaved complex doubles (re,im) in the reference AoS layout
 *     [Nx][Ny][Nzf][3], Nzf = N/2+1, index 3*((i*N+j)*Nzf+k)+d  (Solver/src/solver.c:640-647)
 *   - kx, ky = 0..N/2, -N/2+1..-1 (Nyquist stored as +N/2); kz = 0..N/2 (Solver/src/solver.c:1793-1822)
 *   - forward transform unnormalised, e^{-ik.x}; backward (c2r) unnormalised; the c2r of the
 *     kz=0 and kz=N/2 lines uses only the real parts of those two coefficients (the convention of
 *     numpy.fft.irfftn / FFTW c2r after complex x,y passes)
 *   - real arrays here are unpadded [N][N][N] (z fastest)
 *
 * Reference defects corrected (documented, not reproduced):
 *   D1 FFT plans are built for the true real size {N,N,N}   (Solver/src/solver.c:55,2067-2068)
 *   D2 ApplyDealiasing keeps retained modes unchanged       (Solver/src/solver.c:1741)
 *   D3/D14 measures: correct by default; quirk=1 reproduces the reference's operator-precedence
 *          doubling and conjugate-free helicity              (Solver/src/solver.c:1227-1235)
 *   D16 spectrum bins >= n_spect are dropped (they hold only dealiased, i.e. zero, modes)
 *
 * Parity pins: every function is checked by tests/test_oracle_*.py against numpy golden vectors
 * (tests/golden/), brute-force DFTs, closed forms (Taylor-Green, Beltrami/Shapiro) or invariants.
 * oracle_ic_random_ref is pinned to the glibc rand() stream and to the draw order gcc gives the
 * reference's expression (tests/test_oracle_random_ic.py).
 * D18-D20 (Dormand-Prince tableau, error estimate and step control) are listed at oracle_step_rk5
 * and oracle_dp5_solve in nsg_oracle.c.
 */
#ifndef NSG_ORACLE_H
#define NSG_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* time-stepping schemes */
#define ORACLE_RK4CN 0 /* reference RK4 + Crank-Nicolson viscous factor, Solver/src/solver.c:505-608 */
#define ORACLE_IFRK4 1 /* integrating-factor RK4 (exp(-nu|k|^2p dt)), no reference counterpart */
#define ORACLE_RK5CN 2 /* Dormand-Prince 5th-order weights + Crank-Nicolson factor, Solver/src/solver.c:217-421 (__RK5) */
#define ORACLE_DP5 3   /* same update; the adaptive controller is oracle_dp5_solve (__DPRK5) */

int oracle_n_spect(int N);
void oracle_wavenumbers(int N, int* kx, int* ky, int* kz);
void oracle_set_threads(int nthreads);

/* single-component 3-D transforms: real [N][N][N] <-> complex [N][N][Nzf] (interleaved) */
void oracle_r2c3d(int N, const double* in, double* out);
void oracle_c2r3d(int N, const double* in, double* out);
/* 1-D complex transform of length n (sign -1 forward, +1 backward), unnormalised */
void oracle_fft1d(int n, double* data, int sign);

void oracle_dealias(int N, double* uh);
void oracle_curl(int N, const double* uh, double* wh);
void oracle_rhs(int N, const double* uh, double* nl);
void oracle_forcing(int N, const double* uh, double eps_in, double kf, double* nl);
void oracle_step(int N, double* uh, double nu, int visc_pow, double dt, int scheme,
                 double force_eps, double force_kf);
/* one Dormand-Prince RK step from uh into out (may alias uh); nstage 6 (__RK5) or 7 (__DPRK5,
   returns the embedded error norm, 0 otherwise); quirk 1 uses the reference's A32 = 0.5 (D18) */
double oracle_step_rk5(int N, const double* uh, double* out, double nu, int visc_pow, double dt, int nstage,
                       int quirk, double force_eps, double force_kf);
/* step-size factor min(1.5, max(0.01, 0.8 err^(-1/5))) */
double oracle_dp5_factor(double err);
/* nsteps accepted adaptive steps; dt and t updated in place; 0 or -1 (50 rejections / non-finite) */
int oracle_dp5_solve(int N, double* uh, double nu, int visc_pow, double* dt, double* t, int nsteps,
                     double force_eps, double force_kf, long* rejected, double* last_err);
void oracle_measures(int N, const double* uh, double nu, int visc_pow, int quirk,
                     double* totals5, double* Ek, double* Ensk);
void oracle_realspace(int N, const double* uh, int which, double* out);
double oracle_max_speed_sum(int N, const double* uh);
double oracle_cfl_dt(int N, const double* uh, double cfl);

void oracle_ic_taylor_green(int N, double* uh);
void oracle_ic_shapiro(int N, double nu, double t, double* uh);
void oracle_ic_random_ref(int N, double* uh);
void oracle_ic_k4(int N, double k0, double energy, uint64_t seed, double* uh);

void oracle_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* post-processing increment statistics (stats_oracle.c, PostProcessing/src/stats.c) */
#define ORACLE_NUM_INCR 9
#define ORACLE_N_BINS 1000
void oracle_incr_list(int N, int* r);
/* acc [9][7] = {n, min, max, mean, M2, M3, M4} per increment, updated in the reference's order */
void oracle_incr_moments(int N, const double* u, double* acc);
void oracle_incr_ranges(const double* acc, double* ranges);
long oracle_incr_histogram(int N, const double* u, const double* ranges, double* counts);
void oracle_incr_stats(const double* acc, double* out);

#ifdef __cplusplus
}
#endif
#endif
