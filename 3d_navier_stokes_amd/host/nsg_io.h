/* nsg_io.h -- HDF5 output/restart of the host driver (reference layout, Solver/src/hdf5_funcs.c) */
#ifndef NSG_IO_H
#define NSG_IO_H
#include <hdf5.h>

#include "nsg_run.h"

typedef struct nsg_h5 {
    hid_t main, spect, ctype;
    char dir[1600], main_name[2000], spect_name[2000];
} nsg_h5;

int nsg_h5_open(nsg_h5* h, const nsg_run* r, const char* solver_tag);
int nsg_h5_write_save(nsg_h5* h, int idx, double t, double dt, int N, const double* u_hat, const double* w_hat,
                      const double* u, const double* w, const double* Ek, const double* Ensk, int nspect);
int nsg_h5_final(nsg_h5* h, int N, int nsave, const double* time, const double* E, const double* Ens, const double* Pal,
                 const double* Hel, const double* Diss);
void nsg_h5_close(nsg_h5* h);
int nsg_h5_read_restart(const char* file, int N, double* u_hat, double* t, double* dt, int* idx);

#endif
