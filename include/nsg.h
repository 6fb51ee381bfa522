/*
 * nsg.h -- C ABI of the MI355X (gfx950) pseudo-spectral Navier-Stokes time-step library.
 *
 * One context owns one periodic box [0,2pi)^3 resident in one GPU's HBM.  Plain pointers and
 * sizes only; spectral host arrays use the reference layout: interleaved complex double
 * (re,im) in AoS order [Nx][Ny][Nz/2+1][3], index 3*((i*Ny+j)*(Nz/2+1)+k)+d, FFTW normalisation
 * u_hat = sum_x u(x) e^{-i k.x} (Solver/src/solver.c:640-647, :1573).  Real host arrays are AoS
 * [Nx][Ny][Nz][3] doubles.  All calls return NSG_OK (0) or a negative NSG_ERR_* code; the text of
 * the last error is nsg_last_error(ctx).  A context is not thread-safe; use one per host thread.
 *
 * The library keeps only the modes retained by the reference's 2/3 rule
 * (|k| <= floor(N/3), Solver/src/solver.c:1731-1737): nsg_set_state applies that mask on upload,
 * exactly as the reference applies ApplyDealiasing to every state it creates.
 */
#ifndef NSG_H
#define NSG_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define NSG_OK 0
#define NSG_ERR_INVALID (-1)   /* bad argument / unsupported size */
#define NSG_ERR_OOM (-2)       /* device or host allocation failed */
#define NSG_ERR_DEVICE (-3)    /* HIP runtime error */
#define NSG_ERR_NONFINITE (-4) /* NaN/Inf in the state or its diagnostics */
#define NSG_ERR_BLOWUP (-5)    /* max|u| exceeded nsg_params.max_speed */
#define NSG_ERR_STEP (-6)      /* an adaptive step was rejected 50 times (Solver/src/data_types.h:128) */

#define NSG_SCHEME_RK4CN 0 /* reference RK4 with Crank-Nicolson viscous factor, Solver/src/solver.c:505-608 */
#define NSG_SCHEME_IFRK4 1 /* integrating-factor RK4 */
#define NSG_SCHEME_RK5CN 2 /* reference RK5 (-D__RK5): Dormand-Prince 5th-order weights, Crank-Nicolson
                              factor, fixed dt, Solver/src/solver.c:217-349 (A32 = 9/40, see nsg_step) */
#define NSG_SCHEME_DP5 3   /* reference adaptive Dormand-Prince (-D__DPRK5), Solver/src/solver.c:134-153,
                              217-421; one box per context, FP64 only, cfl must be 0, dt is the
                              first try's step */

#define NSG_FP64 0     /* every FFT and product in FP64 */
#define NSG_FP32BULK 1 /* FFTs, products and stage increments in FP32; state, RK update,
                          exponentials and diagnostics in FP64 */

#define NSG_IC_TAYLOR_GREEN 0 /* Solver/src/solver.c:1553-1574 */
#define NSG_IC_SHAPIRO 1      /* Solver/src/solver.c:1575-1600 with u_z = A cos cos sin(Mz) (divergence-free) */
#define NSG_IC_K4 2           /* random-phase solenoidal E(k) ~ k^4 exp(-2 (k/k0)^2), Philox-4x32-10 */

typedef struct nsg_params {
    int N[3];          /* grid; cubic, N in {8,12,16,24,32,48,64,96,128,256,512,1024} */
    double nu;         /* viscosity (Solver/src/utils.c:65, -v) */
    int visc_pow;      /* 1: nu|k|^2; 2: nu|k|^4 (reference -D__HYPER build, Solver/src/data_types.h:68-71) */
    double dt;         /* time step (-h); initial step when cfl > 0 */
    double t0;         /* start time (-s) */
    int scheme;        /* NSG_SCHEME_* */
    int precision;     /* NSG_FP64 | NSG_FP32BULK */
    double cfl;        /* > 0: dt = cfl * (2pi/N) / max(|u_x|+|u_y|+|u_z|) each step (-c) */
    double force_eps;  /* > 0: constant-power forcing, volume-averaged injection rate */
    double force_kf;   /* forcing band 0 < |k| <= force_kf (at most 7) */
    double max_speed;  /* > 0: nsg_step fails with NSG_ERR_BLOWUP when max|u| exceeds it */
    int device;        /* HIP device ordinal */
} nsg_params;

typedef struct nsg_totals {
    double energy;       /* TotalEnergy      (Solver/src/solver.c:1274) */
    double enstrophy;    /* TotalEnstrophy   (:1271) */
    double palinstrophy; /* TotalPalinstrophy(:1272) */
    double helicity;     /* TotalHelicity    (:1273) */
    double dissipation;  /* EnergyDissipation(:1270) */
} nsg_totals;

typedef struct nsg_ctx nsg_ctx;

/* defaults of GetCMLArgs (Solver/src/utils.c:45-67): N=64^3, dt=1e-3, nu=1, RK4CN, FP64 */
void nsg_default_params(nsg_params* p);
/* replaces AllocateMemory + InitializeFFTWPlans + InitializeSpaceVariables (Solver/src/solver.c:65-77) */
int nsg_create(const nsg_params* p, nsg_ctx** out);
void nsg_destroy(nsg_ctx* ctx);
const char* nsg_last_error(const nsg_ctx* ctx);
/* n_spect = (int)sqrt(3 (N/2)^2) + 1 (Solver/src/solver.c:1384) */
int nsg_n_spect(const nsg_ctx* ctx);

/* state transfer (run_data->u_hat), reference AoS layout, 2/3 mask applied on upload */
int nsg_set_state(nsg_ctx* ctx, const double* u_hat);
int nsg_get_state(nsg_ctx* ctx, double* u_hat);
/* u_hat = dealias(r2c(u)) from a real AoS field (the IC path, Solver/src/solver.c:1573,1630) */
int nsg_set_state_real(nsg_ctx* ctx, const double* u);
/* InitialConditions (Solver/src/solver.c:1537-1648); k0, energy (1/2<|u|^2>_vol), seed for NSG_IC_K4 */
int nsg_init_state(nsg_ctx* ctx, int kind, double k0, double energy, uint64_t seed);

/* NonlinearRHSBatch (Solver/src/solver.c:620-731), plus forcing when enabled:
 * out = dealias(P(r2c(c2r(u_hat) x c2r(i k x u_hat)))/N^6) + f(u_hat) */
int nsg_rhs(nsg_ctx* ctx, const double* u_hat, double* out);
/* nsteps x RK4Step (Solver/src/solver.c:131 -> :505) on the device-resident state; RK5CN: RK5DPStep
 * (:133 -> :217).  DP5: nsteps accepted adaptive steps (the loop of :134-153); each try runs the
 * seven stages, the embedded error err = max_c sqrt(sum |e_c/(1e-7 + 1e-7 max(|u_c|,|u5_c|))|^2 /
 * (N^2 (N/2+1))) with e = 2dt/(2+D) sum (B_i - Bs_i) K_i, is accepted when err <= 1 (time advances
 * by that try's dt), and sets dt <- dt min(1.5, max(0.01, 0.8 err^-1/5)) either way; 50 rejections
 * in a row fail with NSG_ERR_STEP.  Deliberate differences from the reference's code: A32 = 9/40
 * (the reference's 0.5 makes the scheme first order), the error taken between the 5th- and
 * 4th-order solutions from the same start state with |e|^2 and zeroed sums, the accept test
 * err <= 1 (the reference retries exactly those tries), t = sum of accepted dt, and the first
 * step checked too. */
int nsg_step(nsg_ctx* ctx, long nsteps);
/* Dormand-Prince bookkeeping (RK_data->DP_err, DP_fails, Solver/src/data_types.h:233-234): steps
 * accepted, tries rejected, error norm of the last try; NSG_ERR_INVALID for other schemes */
int nsg_dp_stats(nsg_ctx* ctx, long long* accepted, long long* rejected, double* last_err);
/* ComputeSystemMeasurables (Solver/src/solver.c:1142-1276); E_k/Ens_k may be NULL, else n_spect
 * doubles. quirk = 1 reproduces the reference's operator-precedence (D3) and conjugate-free
 * helicity (D14) arithmetic. */
int nsg_measures(nsg_ctx* ctx, nsg_totals* out, double* E_k, double* Ens_k, int n_spect, int quirk);
/* nsg_measures without allocation.  No counterpart in the reference: it is for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same kernels and same results as nsg_measures (bit for bit), but the per-shell partial sums live
 * in a static device array of the library (64 KiB per device, 8192 doubles, held until the process
 * ends; 7 doubles per spectral shell, i.e. n_spect <= 1170, which every supported N satisfies),
 * and the results go to the caller's DEVICE buffer out[5 + 2 n_spect]: the five totals in
 * nsg_totals order, then E_k[n_spect], then Ens_k[n_spect].  Asynchronous on the context's stream
 * (visible after any synchronising call on the context, or hipDeviceSynchronize); no host wait,
 * no non-finite check.  Calls that overlap on one GPU share that array and race: serialise them
 * (contexts on one device with concurrent streams must not call it at the same time). */
int nsg_measures_noalloc(nsg_ctx* ctx, double* out, int n_spect, int quirk);
/* real-space u (which=0) or w (which=1), scaled 1/N^3 (Solver/src/hdf5_funcs.c:221-223), AoS [N][N][N][3] */
int nsg_realspace(nsg_ctx* ctx, int which, double* out);
/* simulation time, current dt, and max(|u_x|+|u_y|+|u_z|) of the last completed step */
int nsg_get_time(nsg_ctx* ctx, double* t, double* dt, double* max_speed);
int nsg_set_time(nsg_ctx* ctx, double t, double dt);

/* Per-pass timing of one step: runs `iters` steps (DP5: tries, not applied; state and time stay)
 * under HIP events on the library's stream and returns the mean duration (ms) of each kernel
 * class: [0] pass1 inverse-x+curl, [1] pass2 inverse-y, [2] pass3 z (c2r, u x w, r2c), [3] pass4
 * forward-y, [4] pass5 forward-x+epilogue, [5] whole step.  bytes[0..4] receive the algorithmic
 * HBM bytes of one launch of each pass, averaged over the stages of a step. */
int nsg_profile_passes(nsg_ctx* ctx, int iters, double* ms, double* bytes);
/* algorithmic HBM bytes of one full step (sum over its 5 x stages pass launches) */
double nsg_step_bytes(const nsg_ctx* ctx);

/* Launch schedules = FFT schedule f x persistent z-pass grid g, numbered s = g * F + f with F the
 * FFT schedules of N.  Every size has FFT schedule 0; N <= 256 have two alternatives (1: the other
 * x/y radices, shorter or longer than schedule 0's; 2: wider column tiles), so F = 3 there and 1
 * elsewhere.  The z grids are 8, 4 and 12 blocks per CU (g = 0, 1, 2; 4, 8, 12 at 128^3 FP64).
 * FFT schedules differ only in rounding; the z grid changes no result bit.
 * nsg_schedule_count(N) = 3 F.  nsg_autotune times `iters` steps
 * of each on the context's own state, restores the state and time afterwards, keeps the fastest and
 * caches the choice per (N, precision, boxes) for the process (use_cache = 0 re-measures).
 * ms_per_step, if not NULL, receives one entry per schedule.
 * No reference counterpart: the reference is planned by FFTW_MEASURE (Solver/src/solver.c:2034-2073). */
int nsg_schedule_count(int N);
int nsg_set_schedule(nsg_ctx* ctx, int sched);
int nsg_autotune(nsg_ctx* ctx, int iters, int use_cache, int* chosen, double* ms_per_step);

/* Ensembles: many independent boxes on one GPU.  The reference integrates one box per process,
 * so a parameter sweep is one solver run per case; an ensemble replaces that loop of SpectralSolve
 * runs (Solver/src/solver.c:50-207).  Boxes sharing N, precision, scheme, visc_pow, forcing and device
 * form one batch (one launch per pass covers the whole batch); nu, dt, t0, cfl and max_speed are
 * per box.  nsg_ensemble_step replays a HIP graph holding one step of every batch, the batches
 * on their own streams; each box evolves exactly as it would alone under the same schedule.
 * Box indices follow the order of `boxes`. */
typedef struct nsg_ensemble nsg_ensemble;
int nsg_ensemble_create(const nsg_params* boxes, int nbox, nsg_ensemble** out);
void nsg_ensemble_destroy(nsg_ensemble* e);
const char* nsg_ensemble_last_error(const nsg_ensemble* e);
int nsg_ensemble_group_count(const nsg_ensemble* e);
int nsg_ensemble_group_of(const nsg_ensemble* e, int box, int* group, int* index);
int nsg_ensemble_n_spect(const nsg_ensemble* e, int box);
int nsg_ensemble_init_state(nsg_ensemble* e, int box, int kind, double k0, double energy, uint64_t seed);
int nsg_ensemble_set_state(nsg_ensemble* e, int box, const double* u_hat);
int nsg_ensemble_get_state(nsg_ensemble* e, int box, double* u_hat);
int nsg_ensemble_measures(nsg_ensemble* e, int box, nsg_totals* out, double* E_k, double* Ens_k, int n_spect, int quirk);
int nsg_ensemble_realspace(nsg_ensemble* e, int box, int which, double* out);
int nsg_ensemble_get_time(nsg_ensemble* e, int box, double* t, double* dt, double* max_speed);
/* nsteps steps of every box: graph replays (nsg_ensemble_step) or one box at a time with direct
 * launches (nsg_ensemble_step_serial, the serial per-box reference) */
int nsg_ensemble_step(nsg_ensemble* e, long nsteps);
int nsg_ensemble_step_serial(nsg_ensemble* e, long nsteps);
/* nsteps steps timed with HIP events on the ensemble's stream (serial = 0: graph; 1: per box) */
int nsg_ensemble_time_steps(nsg_ensemble* e, long nsteps, int serial, double* ms);
int nsg_ensemble_autotune(nsg_ensemble* e, int iters, int use_cache, int* chosen_per_group);
int nsg_ensemble_set_schedule(nsg_ensemble* e, int group, int sched);
double nsg_ensemble_step_bytes(const nsg_ensemble* e);

/* Learned subgrid-stress closure network (no reference counterpart: the reference has no LES).
 * Four periodic 3x3x3 convolutions 9 -> 32 -> 32 -> 32 -> 6 channels with exact GELU between
 * them, on an N^3 grid (N a multiple of 16); BF16 weights and activations, FP32 accumulation on
 * the matrix cores.  Input: the 9 velocity-gradient components g[3 i + j] = d_j u_i of the
 * filtered field; output: the deviatoric subgrid stress (xx, yy, zz, xy, xz, yz).  Host arrays
 * are channels-last float [batch][N^3][C], point index (x N + y) N + z.  Parameters are
 * exchanged as floats in torch conv3d order, layer by layer: W[co][ci][3][3][3] then b[co].
 * nsg_cnn_forward / nsg_cnn_loss_grad take normalised inputs and targets; nsg_cnn_add_sample
 * takes physical ones and divides by the normalisation set with nsg_cnn_set_norm (inputs by
 * sigma_in, stress component c by sigma_out[c]).  Loss: mean squared error over points and the
 * 6 outputs; training: Adam (beta 0.9, 0.999, eps 1e-8) on FP32 master weights. */
typedef struct nsg_cnn nsg_cnn;
long long nsg_cnn_param_count(void);
int nsg_cnn_create(int N, int max_batch, uint64_t seed, int device, nsg_cnn** out);
void nsg_cnn_destroy(nsg_cnn* c);
const char* nsg_cnn_last_error(const nsg_cnn* c);
int nsg_cnn_set_params(nsg_cnn* c, const float* params);
int nsg_cnn_get_params(nsg_cnn* c, float* params);
int nsg_cnn_set_norm(nsg_cnn* c, double sigma_in, const double* sigma_out);
int nsg_cnn_get_norm(const nsg_cnn* c, double* sigma_in, double* sigma_out);
int nsg_cnn_forward(nsg_cnn* c, const float* x, int batch, float* y);
/* loss and its gradient with respect to every parameter (grads may be NULL: loss only) */
int nsg_cnn_loss_grad(nsg_cnn* c, const float* x, const float* target, int batch, double* loss, float* grads);
/* training set resident on the device: one sample = one N^3 field pair */
int nsg_cnn_add_sample(nsg_cnn* c, const float* x, const float* tau);
int nsg_cnn_sample_count(const nsg_cnn* c);
/* `steps` Adam steps on random batches drawn from samples [first, first + count) (indices from
 * `seed`); loss_per_step may be NULL */
int nsg_cnn_train(nsg_cnn* c, int first, int count, int steps, int batch, double lr, uint64_t seed, double* loss_per_step);
/* mean loss over samples [first, first + count) */
int nsg_cnn_eval(nsg_cnn* c, int first, int count, double* loss);
/* mean time (ms, HIP events) of one inference pass over `batch` fields */
int nsg_cnn_time_forward(nsg_cnn* c, int batch, int iters, double* ms);

/* LES with the learned closure (no reference counterpart).  Filter: Gaussian
 * G(k) = exp(-k^2 Delta^2 / 24), Delta = pi / kc, times the LES grid's 2/3 cut |k| <= N_les / 3.
 * nsg_sgs_sample: from a DNS context's state, writes the filtered state into the (coarser) LES
 * context and, when the pointers are not NULL, the model input x = grad(u_filtered)
 * [N_les^3][9] (x[3 i + j] = d_j u_i) and the target: the deviatoric part of
 * tau_ij = F(u_i u_j) - T(ub_i ub_j) [N_les^3][6] (F: the filter, T: the sharp cut).
 * nsg_attach_closure: every later nsg_step of the (single-box) context adds
 * -P div(tau_model(grad u)) to the nonlinear term of each RK stage, evaluated from the state at
 * the start of the step; NULL detaches.  The network must live as long as it is attached.
 * nsg_closure_term: that term for the current state, reference AoS half-spectrum layout. */
int nsg_sgs_sample(nsg_ctx* dns, nsg_ctx* les, double kc, float* x_out, float* tau_out);
/* spectral resampling between grids (truncation to, or zero-padding from, the coarser grid's
 * 2/3 sphere), FFTW normalisation rescaled by (N_dst / N_src)^3; time copied, dt scaled by
 * N_src / N_dst (the CFL ratio) */
int nsg_resample(nsg_ctx* src, nsg_ctx* dst);
int nsg_attach_closure(nsg_ctx* ctx, nsg_cnn* net);
int nsg_closure_term(nsg_ctx* ctx, double* out);
/* per-shell gains of the attached closure term (no reference counterpart): the term's modes on the
 * shell b = round(|k|) (the energy spectrum's shells) are multiplied by gain[b], b = 0..kmax
 * (n = kmax + 1); NULL removes them.  They stay until removed or the closure is detached. */
int nsg_closure_set_gain(nsg_ctx* ctx, const double* gain, int n);

/* ---------------------------------------------------------------------------------------------
 * Velocity-increment statistics (the reference's post-processing, PostProcessing/src/stats.c).
 * Longitudinal increments u_x(i+r,j,k) - u_x(i,j,k), u_y(i,j+r,k) - u_y(i,j,k),
 * u_z(i,j,k+r) - u_z(i,j,k), shifted index inside the box (no periodic wrap), all three into one
 * accumulator per separation r in {1,2,4,8,16,32,64,128,N/2} (AllocateStatsObjects, :404-413;
 * NUM_INCR 9, N_BINS 1000, PostProcessing/src/data_types.h:49-51).  Real fields are AoS
 * [N][N][N][3] doubles (the post-processing's run_data->u).
 * ------------------------------------------------------------------------------------------- */
#define NSG_INCR_NUM 9
#define NSG_INCR_BINS 1000
typedef struct nsg_incr nsg_incr;
/* gsl_rstat_alloc / gsl_histogram_alloc per separation (stats.c:404-423) */
int nsg_incr_create(int N, int device, nsg_incr** out);
void nsg_incr_destroy(nsg_incr* s);
const char* nsg_incr_last_error(const nsg_incr* s);
/* the separations r[NSG_INCR_NUM] */
int nsg_incr_increments(const nsg_incr* s, int* r);
/* Precompute's pass over one snapshot (stats.c:62-112): count, min, max, mean and central moments
 * of every separation's increments, accumulated over calls.  _state: the snapshot is the
 * context's real-space velocity (1/N^3-normalised, as nsg_realspace) times `scale`; scale = N^3
 * gives the post-processing's unnormalised c2r (PostProcessing/src/hdf5_func.c:306) */
int nsg_incr_precompute(nsg_incr* s, const double* u);
/* (the _state calls leave the field in the context's scratch, which its next call reuses) */
int nsg_incr_precompute_state(nsg_incr* s, nsg_ctx* ctx, double scale);
/* end of Precompute (stats.c:118-132): bin ranges [min - 0.05, max + 0.05], uniform as
 * gsl_histogram_set_ranges_uniform; clears the counts.  NSG_ERR_NONFINITE for NaN/Inf input */
int nsg_incr_set_ranges(nsg_incr* s);
/* ComputeStats (stats.c:153-250): histogram counts of one snapshot, accumulated over calls.  An
 * increment outside the ranges (where the reference exits, :191-195) gives NSG_ERR_INVALID; the
 * in-range increments of that snapshot stay counted */
int nsg_incr_histogram(nsg_incr* s, const double* u);
int nsg_incr_histogram_state(nsg_incr* s, nsg_ctx* ctx, double scale);
/* the arrays WriteStatsData writes (stats.c:316-371), any may be NULL: stats [9][6] = min, max,
 * mean, sd, skew, kurtosis (gsl_rstat_*: sd with n - 1, skew and kurtosis over that sd, 0 for an
 * empty set), ranges [9][1001], counts [9][1000] */
int nsg_incr_results(nsg_incr* s, double* stats, double* ranges, double* counts);
/* zero the accumulators and forget the ranges (no reference counterpart) */
int nsg_incr_reset(nsg_incr* s);
/* mean time (ms, HIP events) of one moments (hist 0) or histogram (hist 1) pass over the last
 * host snapshot; the passes accumulate (no reference counterpart) */
int nsg_incr_time_pass(nsg_incr* s, int hist, int iters, double* ms);

#ifdef __cplusplus
}
#endif
#endif
