/*
 * nsg_main.c -- host driver: the reference's main + SpectralSolve (Solver/src/main.c:34-118,
 * Solver/src/solver.c:50-208) on one GPU through the C ABI of libnsg.so.
 *
 * Loop semantics kept from SpectralSolve: num_t_steps = (T - t0)/dt and
 * num_print_steps = num_t_steps/SAVE_EVERY + 1 (InitializeIntegrationVariables,
 * Solver/src/solver.c:1338,1355); t += dt; iters = 1; while (t <= T) { step; save when
 * iters % SAVE_EVERY == 0 } with t = t0 + iters*dt (the reference drops t0, D6).  With --adaptive
 * the step size comes from the device CFL reduction and t is the device's accumulated time.
 * The state never leaves the GPU between saves; fixed-dt steps between two saves go to the library
 * as one nsg_step call.
 */
#define _GNU_SOURCE /* sincos */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "nsg_io.h"
#include "nsg_run.h"

#define CHECK(call)                                                                              \
    do {                                                                                         \
        int rc_ = (call);                                                                        \
        if (rc_ != NSG_OK) {                                                                     \
            fprintf(stderr, "[ERROR] %s -> %d: %s\n", #call, rc_, nsg_last_error(ctx));          \
            nsg_destroy(ctx);                                                                    \
            return 1;                                                                            \
        }                                                                                        \
    } while (0)

/* RANDOM initial condition of the reference (Solver/src/solver.c:1551,1605-1624): srand(123456789),
 * per mode and component amplitude U[0,1] then phase U[0,1] -- the order gcc evaluates the two
 * rand() calls of the reference's expression in (left factor first) -- and cexp of the phase,
 * which glibc computes with sincos; then the 2/3 mask, which nsg_set_state applies. */
static void reference_random_ic(int N, double* uh) {
    const long n = (long)N * N * (N / 2 + 1) * 3;
    srand(123456789);
    for (long i = 0; i < n; ++i) {
        double amp = (double)rand() / (double)RAND_MAX;
        double ph = (double)rand() / (double)RAND_MAX;
        double sn, cs;
        sincos(ph * 2.0 * M_PI, &sn, &cs);
        uh[2 * i] = amp * cs;
        uh[2 * i + 1] = amp * sn;
    }
}

/* w_hat = i k x u_hat on the host copy (reference AoS layout) */
static void curl_host(int N, const double* uh, double* wh) {
    const int Nzf = N / 2 + 1;
    for (int i = 0; i < N; ++i) {
        const double kx = (i <= N / 2) ? i : i - N;
        for (int j = 0; j < N; ++j) {
            const double ky = (j <= N / 2) ? j : j - N;
            for (int k = 0; k < Nzf; ++k) {
                const double kz = k;
                const long b = 6 * (((long)i * N + j) * Nzf + k);
                const double* u = uh + b;
                double* w = wh + b;
                const double ar = ky * u[4] - kz * u[2], ai = ky * u[5] - kz * u[3];
                const double br = kz * u[0] - kx * u[4], bi = kz * u[1] - kx * u[5];
                const double cr = kx * u[2] - ky * u[0], ci = kx * u[3] - ky * u[1];
                w[0] = -ai; w[1] = ar; w[2] = -bi; w[3] = br; w[4] = -ci; w[5] = cr;
            }
        }
    }
}

typedef struct {
    double *time, *E, *Ens, *Pal, *Hel, *Diss;
    long cap, n;
} series;

static int record(nsg_ctx* ctx, const nsg_run* r, nsg_h5* h, series* s, int save_idx, double t, double dt, int N,
                  double* uh, double* wh, double* ur, double* wr, double* ek, double* ens, int nspect, long iters,
                  long nsteps, double T) {
    nsg_totals m;
    int rc = nsg_measures(ctx, &m, ek, ens, nspect, r->quirk);
    if (rc != NSG_OK) {
        fprintf(stderr, "[ERROR] measures: %s\n", nsg_last_error(ctx));
        return rc;
    }
    if (s->n < s->cap) {
        s->time[s->n] = t; s->E[s->n] = m.energy; s->Ens[s->n] = m.enstrophy; s->Pal[s->n] = m.palinstrophy;
        s->Hel[s->n] = m.helicity; s->Diss[s->n] = m.dissipation;
        s->n++;
    } else {
        printf("\n[WARNING] --- Unable to write system measures at Indx: [%d] t: [%lf]\n", save_idx, t);
    }
    if ((rc = nsg_get_state(ctx, uh)) != NSG_OK) return rc;
    curl_host(N, uh, wh);
    if (r->realspace) {
        if ((rc = nsg_realspace(ctx, 0, ur)) != NSG_OK) return rc;
        if ((rc = nsg_realspace(ctx, 1, wr)) != NSG_OK) return rc;
    }
    if (nsg_h5_write_save(h, save_idx, t, dt, N, uh, wh, r->realspace ? ur : NULL, r->realspace ? wr : NULL,
                          r->spectra ? ek : NULL, r->spectra ? ens : NULL, nspect) != 0) {
        fprintf(stderr, "[ERROR] HDF5 write of /Iter_%05d failed\n", save_idx);
        return -10;
    }
    if (!r->quiet) {
        double tt, ddt, mx;
        nsg_get_time(ctx, &tt, &ddt, &mx);
        printf("Iter: %ld/%ld\tt: %1.6lf/%1.3lf\tdt: %g\tMax Vel: %g\tKE: %g\tENS: %g\tPAL: %g\tHEL: %g\tDISS: %g\n", iters,
               nsteps, t, T, dt, mx, m.energy, m.enstrophy, m.palinstrophy, m.helicity, m.dissipation);
        fflush(stdout);
    }
    return NSG_OK;
}

int main(int argc, char** argv) {
    nsg_run r;
    int prc = nsg_parse_args(argc, argv, &r);
    if (prc) return prc == 2 ? 0 : 1;
    if (r.N[0] != r.N[1] || r.N[1] != r.N[2]) {
        fprintf(stderr, "[ERROR] only cubic grids are supported (got %d x %d x %d)\n", r.N[0], r.N[1], r.N[2]);
        return 1;
    }
    const int N = r.N[0];
    clock_t begin = clock();

    nsg_params p;
    nsg_default_params(&p);
    p.N[0] = p.N[1] = p.N[2] = N;
    p.nu = r.nu;
    p.visc_pow = r.visc_pow;
    p.dt = r.dt;
    p.t0 = r.t0;
    p.scheme = r.scheme;
    p.precision = r.precision;
    p.cfl = r.adaptive ? r.cfl : 0.0;
    p.force_eps = strcmp(r.forcing, "CONST_POWER") == 0 ? r.force_eps : 0.0;
    p.force_kf = r.force_kf;
    p.max_speed = r.max_speed;
    p.device = r.device;
    nsg_ctx* ctx = NULL;
    CHECK(nsg_create(&p, &ctx));
    const int nspect = nsg_n_spect(ctx);

    const size_t nspec = (size_t)N * N * (N / 2 + 1) * 3 * 2;
    const size_t nreal = (size_t)N * N * N * 3;
    double* uh = calloc(nspec, sizeof(double));
    double* wh = calloc(nspec, sizeof(double));
    double* ur = r.realspace ? malloc(nreal * sizeof(double)) : NULL;
    double* wr = r.realspace ? malloc(nreal * sizeof(double)) : NULL;
    double* ek = calloc((size_t)nspect, sizeof(double));
    double* ens = calloc((size_t)nspect, sizeof(double));
    if (!uh || !wh || (r.realspace && (!ur || !wr)) || !ek || !ens) {
        fprintf(stderr, "[ERROR] host allocation failed\n");
        nsg_destroy(ctx);
        return 1;
    }

    /* the Dormand-Prince scheme chooses dt itself: step by step, like the CFL-adaptive loop */
    const int var_dt = r.adaptive || r.scheme == NSG_SCHEME_DP5;
    /* initial condition (InitialConditions, Solver/src/solver.c:1537-1648) or restart */
    double t0 = r.t0, dt = r.dt;
    int save_idx = 0;
    if (strcmp(r.input_file, "NONE") != 0) {
        if (nsg_h5_read_restart(r.input_file, N, uh, &t0, &dt, &save_idx) != 0) {
            fprintf(stderr, "[ERROR] cannot read a restart state from %s\n", r.input_file);
            nsg_destroy(ctx);
            return 1;
        }
        /* fixed-dt runs step with -h (the loop counts steps of r.dt); a variable-dt run continues
         * from the step size the file was saved with */
        if (!var_dt) dt = r.dt;
        CHECK(nsg_set_state(ctx, uh));
        CHECK(nsg_set_time(ctx, t0, dt));
        r.t0 = t0;
        printf("Restart from %s /Iter_%05d at t = %g, dt = %g\n", r.input_file, save_idx, t0, dt);
        save_idx = 0;
    } else if (!strcmp(r.u0, "TAYLOR_GREEN")) {
        CHECK(nsg_init_state(ctx, NSG_IC_TAYLOR_GREEN, 0, 0, 0));
    } else if (!strcmp(r.u0, "SHAPIRO")) {
        CHECK(nsg_init_state(ctx, NSG_IC_SHAPIRO, 0, 0, 0));
    } else if (!strcmp(r.u0, "K4")) {
        CHECK(nsg_init_state(ctx, NSG_IC_K4, r.k0, r.energy, r.seed));
    } else if (!strcmp(r.u0, "TESTING")) {
        CHECK(nsg_set_state(ctx, uh)); /* the reference's TESTING branch is empty: zero field */
    } else {
        printf("\n[WARNING] --- No initial conditions specified\n---> Using random initial conditions...\n");
        reference_random_ic(N, uh);
        CHECK(nsg_set_state(ctx, uh));
    }

    /* integration counters (InitializeIntegrationVariables, Solver/src/solver.c:1314-1363); with
     * --transients the first TRANS_FRAC of the steps are not saved and the initial condition is
     * not written (:1339-1348, Solver/src/hdf5_funcs.c:117) */
    const long num_t_steps = (long)((r.T - r.t0) / r.dt);
    const long trans_steps = r.transients ? (long)(0.2 * (double)num_t_steps) : 0;
    long num_print_steps;
    if (r.transients)
        num_print_steps = (num_t_steps >= r.save_every) ? (num_t_steps - trans_steps) / r.save_every
                                                        : num_t_steps - trans_steps;
    else
        num_print_steps = (num_t_steps >= r.save_every) ? num_t_steps / r.save_every + 1 : num_t_steps + 1;
    if (var_dt) num_print_steps *= 2; /* room for extra saves, as __ADAPTIVE_STEP does (:1375-1379) */
    if (!r.quiet) {
        if (r.transients)
            printf("Total Iters: %ld\t Saving Iters: %ld\t Transient Steps: %ld\n", num_t_steps, num_print_steps,
                   trans_steps);
        else
            printf("Total Iters: %ld\t Saving Iters: %ld\n", num_t_steps, num_print_steps);
    }
    series s = {0};
    s.cap = num_print_steps;
    s.time = calloc((size_t)s.cap, sizeof(double)); s.E = calloc((size_t)s.cap, sizeof(double));
    s.Ens = calloc((size_t)s.cap, sizeof(double)); s.Pal = calloc((size_t)s.cap, sizeof(double));
    s.Hel = calloc((size_t)s.cap, sizeof(double)); s.Diss = calloc((size_t)s.cap, sizeof(double));

    nsg_h5 h;
    /* solver tag of the output names (Solver/src/hdf5_funcs.c:372-380) */
    const char* tag = r.scheme == NSG_SCHEME_RK5CN ? "RK5" : r.scheme == NSG_SCHEME_DP5 ? "DP5" : "RK4";
    if (nsg_h5_open(&h, &r, tag) != 0) { nsg_destroy(ctx); return 1; }
    if (!r.quiet) printf("\nMain Output File: %s\n\n", h.main_name);

    int rc = NSG_OK;
    if (!r.transients)
        rc = record(ctx, &r, &h, &s, save_idx++, r.t0, dt, N, uh, wh, ur, wr, ek, ens, nspect, 0, num_t_steps, r.T);
    /* time loop */
    double t = r.t0 + dt;
    long iters = 1, pending = 0;
    while (rc == NSG_OK && t <= r.T) {
        const int save = iters > trans_steps && (iters % r.save_every) == 0;
        if (r.transients && iters == trans_steps && !r.quiet) printf("\n\n...Transient Iterations Complete!\n\n");
        if (var_dt) {
            rc = nsg_step(ctx, 1);
            if (rc == NSG_OK) nsg_get_time(ctx, &t, &dt, NULL);
        } else {
            pending++;
            const double tn = r.t0 + (double)(iters + 1) * r.dt;
            if (save || !(tn <= r.T)) {
                rc = nsg_step(ctx, pending);
                pending = 0;
            }
        }
        if (rc != NSG_OK) {
            fprintf(stderr, "[ERROR] nsg_step at iteration %ld: %s\n", iters, nsg_last_error(ctx));
            break;
        }
        if (save) rc = record(ctx, &r, &h, &s, save_idx++, t, dt, N, uh, wh, ur, wr, ek, ens, nspect, iters, num_t_steps, r.T);
        iters++;
        if (!var_dt) t = r.t0 + (double)iters * r.dt;
        else {
            double tn, dtn;
            nsg_get_time(ctx, &tn, &dtn, NULL);
            t = tn + dtn;
        }
    }
    /* FinalWriteAndCloseOutputFile (Solver/src/hdf5_funcs.c:1028-1165) */
    if (nsg_h5_final(&h, N, (int)s.n, s.time, s.E, s.Ens, s.Pal, s.Hel, s.Diss) != 0) rc = rc ? rc : -10;
    nsg_h5_close(&h);
    nsg_destroy(ctx);
    free(uh); free(wh); free(ur); free(wr); free(ek); free(ens);
    free(s.time); free(s.E); free(s.Ens); free(s.Pal); free(s.Hel); free(s.Diss);
    const double spent = (double)(clock() - begin) / CLOCKS_PER_SEC;
    printf("\n\nTotal Execution Time: [%5.10lf] --> %d hrs : %d mins : %d secs\n\n", spent, (int)spent / 3600,
           ((int)spent % 3600) / 60, (int)spent % 60);
    return rc == NSG_OK ? 0 : 1;
}
