/* Test harness (CPU only): drives the host driver's HDF5 writer and restart reader with synthetic
 * data so the output layout can be checked without a GPU.  Built by tests/test_host_io.py. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "nsg_io.h"

int nsg_parse_args(int argc, char** argv, nsg_run* r);

int main(int argc, char** argv) {
    nsg_run r;
    if (nsg_parse_args(argc, argv, &r)) return 3;
    const int N = r.N[0], ns = (int)(1.7320508075688772 * N / 2) + 1;
    const size_t nk = (size_t)N * N * (N / 2 + 1) * 3, nr = (size_t)N * N * N * 3;
    double* uh = malloc(nk * 2 * sizeof(double));
    double* ur = malloc(nr * sizeof(double));
    double* ek = malloc(ns * sizeof(double));
    for (size_t i = 0; i < nk * 2; ++i) uh[i] = (double)i * 0.5;
    for (size_t i = 0; i < nr; ++i) ur[i] = -(double)i;
    for (int i = 0; i < ns; ++i) ek[i] = i + 0.25;
    nsg_h5 h;
    if (nsg_h5_open(&h, &r, "RK4")) return 4;
    double tm[3] = {0.0, 0.5, 1.0}, e[3] = {1, 2, 3};
    for (int s = 0; s < 3; ++s)
        if (nsg_h5_write_save(&h, s, tm[s], 0.125, N, uh, uh, ur, ur, ek, ek, ns)) return 5;
    if (nsg_h5_final(&h, N, 3, tm, e, e, e, e, e)) return 6;
    nsg_h5_close(&h);
    double t, dt;
    int idx;
    double* back = calloc(nk * 2, sizeof(double));
    if (nsg_h5_read_restart(h.main_name, N, back, &t, &dt, &idx)) return 7;
    if (idx != 2 || t != 1.0 || dt != 0.125 || memcmp(back, uh, nk * 2 * sizeof(double))) return 8;
    printf("%s\n%s\n", h.main_name, h.spect_name);
    return 0;
}
