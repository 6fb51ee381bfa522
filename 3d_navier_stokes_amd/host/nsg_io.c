/*
 * nsg_io.c -- HDF5 output in the reference layout, written serially from one process after the
 * device-to-host copy (the reference writes collectively through MPI-IO,
 * Solver/src/hdf5_funcs.c:67-68,861-862,959-960).
 *
 *   <output_dir>/SIM_DATA_NAVIER_RK4_FULL_N[Nx,Ny]_T[t0-T]_NU[nu]_CFL[cfl]_u0[IC]_TAG[tag]/
 *       Main_HDF_Data.h5     /Iter_%05d {u_hat, w_hat: compound{r,i} [Nx][Ny][Nz/2+1][3];
 *                                        u, w: f64 [Nx][Ny][Nz][3]} + attributes TimeValue, TimeStep
 *                            root: kx, ky, kz (int), x, y, z (f64), Time, TotalEnergy,
 *                                  TotalEnstrophy, TotalPalinstrophy, TotalHelicity, EnergyDissipation
 *       Spectra_HDF_Data.h5  /Iter_%05d {EnergySpectrum, EnstrophySpectrum: f64 [n_spect]}
 * (Solver/src/hdf5_funcs.c:33-300 initial write, :446-477 names, :492-760 per-save groups,
 *  :1028-1165 final datasets, :1234-1280 groups and attributes, :1300-1333 complex type).
 * w_hat is the true curl of u_hat (the reference writes an array it never fills, D4).
 */
#include "nsg_io.h"

#include <errno.h>
#include <hdf5.h>
#include <hdf5_hl.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>

typedef struct { double r, i; } h5_complex;

static hid_t complex_type(void) {
    hid_t t = H5Tcreate(H5T_COMPOUND, sizeof(h5_complex));
    H5Tinsert(t, "r", HOFFSET(h5_complex, r), H5T_NATIVE_DOUBLE);
    H5Tinsert(t, "i", HOFFSET(h5_complex, i), H5T_NATIVE_DOUBLE);
    return t;
}

int nsg_h5_open(nsg_h5* h, const nsg_run* r, const char* solver_tag) {
    memset(h, 0, sizeof *h);
    h->main = h->spect = -1;
    char data[1024];
    if (r->file_only) {
        snprintf(data, sizeof data, "_SIM[NAV-%s-FULL]_N[%d,%d]_T[%d-%d]_NU[%1.10lf]_CFL[%1.2lf]_u0[%s].h5", solver_tag, r->N[0],
                 r->N[1], (int)r->t0, (int)r->T, r->nu, r->cfl, r->u0);
        snprintf(h->main_name, sizeof h->main_name, "%sMain_HDF_Data%s", r->output_dir, data);
        snprintf(h->spect_name, sizeof h->spect_name, "%sSpectra_HDF_Data%s", r->output_dir, data);
    } else {
        snprintf(data, sizeof data, "SIM_DATA_NAVIER_%s_FULL_N[%d,%d]_T[%d-%d]_NU[%1.10lf]_CFL[%1.2lf]_u0[%s]_TAG[%s]/", solver_tag,
                 r->N[0], r->N[1], (int)r->t0, (int)r->T, r->nu, r->cfl, r->u0, r->output_tag);
        snprintf(h->dir, sizeof h->dir, "%s%s", r->output_dir, data);
        struct stat st;
        if (stat(h->dir, &st) == -1 && mkdir(h->dir, 0700) == -1) {
            fprintf(stderr, "[ERROR] Unable to create folder for output files [%s]: %s\n", h->dir, strerror(errno));
            return -1;
        }
        snprintf(h->main_name, sizeof h->main_name, "%sMain_HDF_Data.h5", h->dir);
        snprintf(h->spect_name, sizeof h->spect_name, "%sSpectra_HDF_Data.h5", h->dir);
    }
    h->ctype = complex_type();
    h->main = H5Fcreate(h->main_name, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (h->main < 0) {
        fprintf(stderr, "[ERROR] Could not create main HDF5 output file at: %s\n", h->main_name);
        return -1;
    }
    if (r->spectra) {
        h->spect = H5Fcreate(h->spect_name, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
        if (h->spect < 0) {
            fprintf(stderr, "[ERROR] Could not create HDF5 spectra output file at: %s\n", h->spect_name);
            return -1;
        }
    }
    return 0;
}

static hid_t group(hid_t f, const char* name, double t, double dt) {
    if (H5Lexists(f, name, H5P_DEFAULT) > 0) return H5Gopen2(f, name, H5P_DEFAULT);
    hid_t g = H5Gcreate2(f, name, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    if (g < 0) return g;
    hsize_t one = 1;
    hid_t sp = H5Screate_simple(1, &one, NULL);
    hid_t a = H5Acreate2(g, "TimeValue", H5T_NATIVE_DOUBLE, sp, H5P_DEFAULT, H5P_DEFAULT);
    H5Awrite(a, H5T_NATIVE_DOUBLE, &t);
    H5Aclose(a);
    a = H5Acreate2(g, "TimeStep", H5T_NATIVE_DOUBLE, sp, H5P_DEFAULT, H5P_DEFAULT);
    H5Awrite(a, H5T_NATIVE_DOUBLE, &dt);
    H5Aclose(a);
    H5Sclose(sp);
    return g;
}

static int write4(hid_t g, const char* name, hid_t ftype, hid_t mtype, const hsize_t dims[4], const void* data) {
    hid_t sp = H5Screate_simple(4, dims, NULL);
    hid_t d = H5Dcreate2(g, name, ftype, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    int rc = (d < 0) ? -1 : (H5Dwrite(d, mtype, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0 ? -1 : 0);
    if (d >= 0) H5Dclose(d);
    H5Sclose(sp);
    return rc;
}

int nsg_h5_write_save(nsg_h5* h, int idx, double t, double dt, int N, const double* u_hat, const double* w_hat,
                      const double* u, const double* w, const double* Ek, const double* Ensk, int nspect) {
    char name[64];
    snprintf(name, sizeof name, "/Iter_%05d", idx);
    hid_t g = group(h->main, name, t, dt);
    if (g < 0) return -1;
    const hsize_t dk[4] = {(hsize_t)N, (hsize_t)N, (hsize_t)(N / 2 + 1), 3};
    const hsize_t dr[4] = {(hsize_t)N, (hsize_t)N, (hsize_t)N, 3};
    int rc = 0;
    if (w) rc |= write4(g, "w", H5T_NATIVE_DOUBLE, H5T_NATIVE_DOUBLE, dr, w);
    if (w_hat) rc |= write4(g, "w_hat", h->ctype, h->ctype, dk, w_hat);
    if (u_hat) rc |= write4(g, "u_hat", h->ctype, h->ctype, dk, u_hat);
    if (u) rc |= write4(g, "u", H5T_NATIVE_DOUBLE, H5T_NATIVE_DOUBLE, dr, u);
    H5Gclose(g);
    if (h->spect >= 0 && Ek && Ensk) {
        hid_t s = group(h->spect, name, t, dt);
        hsize_t n = (hsize_t)nspect;
        rc |= H5LTmake_dataset(s, "EnstrophySpectrum", 1, &n, H5T_NATIVE_DOUBLE, Ensk) < 0;
        rc |= H5LTmake_dataset(s, "EnergySpectrum", 1, &n, H5T_NATIVE_DOUBLE, Ek) < 0;
        H5Gclose(s);
        H5Fflush(h->spect, H5F_SCOPE_GLOBAL);
    }
    H5Fflush(h->main, H5F_SCOPE_GLOBAL);
    return rc ? -1 : 0;
}

int nsg_h5_final(nsg_h5* h, int N, int nsave, const double* time, const double* E, const double* Ens, const double* Pal,
                 const double* Hel, const double* Diss) {
    int* k = malloc(sizeof(int) * (size_t)N);
    double* x = malloc(sizeof(double) * (size_t)N);
    for (int i = 0; i < N; ++i) {
        k[i] = (i <= N / 2) ? i : i - N;
        x[i] = 2.0 * M_PI * i / N;
    }
    int* kz = malloc(sizeof(int) * (size_t)(N / 2 + 1));
    for (int i = 0; i < N / 2 + 1; ++i) kz[i] = i;
    hsize_t n = (hsize_t)N, nz = (hsize_t)(N / 2 + 1), ns = (hsize_t)nsave;
    int rc = 0;
    rc |= H5LTmake_dataset(h->main, "kx", 1, &n, H5T_NATIVE_INT, k) < 0;
    rc |= H5LTmake_dataset(h->main, "ky", 1, &n, H5T_NATIVE_INT, k) < 0;
    rc |= H5LTmake_dataset(h->main, "kz", 1, &nz, H5T_NATIVE_INT, kz) < 0;
    rc |= H5LTmake_dataset(h->main, "x", 1, &n, H5T_NATIVE_DOUBLE, x) < 0;
    rc |= H5LTmake_dataset(h->main, "y", 1, &n, H5T_NATIVE_DOUBLE, x) < 0;
    rc |= H5LTmake_dataset(h->main, "z", 1, &n, H5T_NATIVE_DOUBLE, x) < 0;
    rc |= H5LTmake_dataset(h->main, "Time", 1, &ns, H5T_NATIVE_DOUBLE, time) < 0;
    rc |= H5LTmake_dataset(h->main, "TotalEnergy", 1, &ns, H5T_NATIVE_DOUBLE, E) < 0;
    rc |= H5LTmake_dataset(h->main, "TotalEnstrophy", 1, &ns, H5T_NATIVE_DOUBLE, Ens) < 0;
    rc |= H5LTmake_dataset(h->main, "TotalPalinstrophy", 1, &ns, H5T_NATIVE_DOUBLE, Pal) < 0;
    rc |= H5LTmake_dataset(h->main, "TotalHelicity", 1, &ns, H5T_NATIVE_DOUBLE, Hel) < 0;
    rc |= H5LTmake_dataset(h->main, "EnergyDissipation", 1, &ns, H5T_NATIVE_DOUBLE, Diss) < 0;
    free(k);
    free(kz);
    free(x);
    return rc ? -1 : 0;
}

void nsg_h5_close(nsg_h5* h) {
    if (h->main >= 0) H5Fclose(h->main);
    if (h->spect >= 0) H5Fclose(h->spect);
    if (h->ctype > 0) H5Tclose(h->ctype);
    h->main = h->spect = -1;
}

/* restart: the /Iter_* group with the largest index, its u_hat and TimeValue/TimeStep */
typedef struct { int best; } iter_scan;
static herr_t scan_cb(hid_t g, const char* name, const H5L_info_t* info, void* op) {
    (void)g;
    (void)info;
    int v;
    if (sscanf(name, "Iter_%d", &v) == 1 && v > ((iter_scan*)op)->best) ((iter_scan*)op)->best = v;
    return 0;
}

int nsg_h5_read_restart(const char* file, int N, double* u_hat, double* t, double* dt, int* idx) {
    hid_t f = H5Fopen(file, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) return -1;
    iter_scan sc = {-1};
    H5Literate(f, H5_INDEX_NAME, H5_ITER_INC, NULL, scan_cb, &sc);
    if (sc.best < 0) { H5Fclose(f); return -1; }
    char name[64];
    snprintf(name, sizeof name, "/Iter_%05d", sc.best);
    hid_t g = H5Gopen2(f, name, H5P_DEFAULT);
    hid_t a = H5Aopen(g, "TimeValue", H5P_DEFAULT);
    H5Aread(a, H5T_NATIVE_DOUBLE, t);
    H5Aclose(a);
    a = H5Aopen(g, "TimeStep", H5P_DEFAULT);
    H5Aread(a, H5T_NATIVE_DOUBLE, dt);
    H5Aclose(a);
    hid_t d = H5Dopen2(g, "u_hat", H5P_DEFAULT);
    int rc = -1;
    if (d >= 0) {
        hid_t sp = H5Dget_space(d);
        hsize_t dims[4] = {0, 0, 0, 0};
        H5Sget_simple_extent_dims(sp, dims, NULL);
        if (dims[0] == (hsize_t)N && dims[1] == (hsize_t)N && dims[2] == (hsize_t)(N / 2 + 1) && dims[3] == 3) {
            hid_t ct = complex_type();
            rc = H5Dread(d, ct, H5S_ALL, H5S_ALL, H5P_DEFAULT, u_hat) < 0 ? -1 : 0;
            H5Tclose(ct);
        }
        H5Sclose(sp);
        H5Dclose(d);
    }
    H5Gclose(g);
    H5Fclose(f);
    *idx = sc.best;
    return rc;
}
