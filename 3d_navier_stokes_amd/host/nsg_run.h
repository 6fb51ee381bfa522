/*
 * nsg_run.h -- run configuration of the host driver (the reference's sys_vars / file_info,
 * Solver/src/data_types.h:147-256), filled from the command line by nsg_parse_args
 * (mirror of GetCMLArgs, Solver/src/utils.c:34-237).
 */
#ifndef NSG_RUN_H
#define NSG_RUN_H
#include <stdint.h>

#include "../../include/nsg.h"

typedef struct nsg_run {
    /* reference flags (same letters and defaults, Solver/src/utils.c:45-67,73) */
    char output_dir[512]; /* -o (a second -o selects file-only output) */
    int file_only;
    char output_tag[64]; /* -t */
    int N[3];            /* -n (repeated, one per dimension) */
    double t0, T, dt;    /* -s, -e, -h */
    double cfl;          /* -c (reference default sqrt(3); adaptive stepping only with --adaptive) */
    double nu;           /* -v */
    double ekman;        /* -d (accepted, unused as in the reference) */
    char u0[64];         /* -i TAYLOR_GREEN | SHAPIRO | RANDOM | TESTING | K4 (default "TG_VORT") */
    int save_every;      /* -p */
    char forcing[64];    /* -f NONE | ZERO | CONST_POWER */
    char input_file[512];/* -z restart file (Main_HDF_Data.h5 written by this driver) */
    /* additions (long options) */
    int visc_pow;        /* --visc-pow 1|2 */
    int scheme;          /* --scheme rk4cn|ifrk4|rk5cn|dp5 */
    int precision;       /* --precision fp64|fp32bulk */
    int adaptive;        /* --adaptive : CFL step with the -c constant */
    double force_eps;    /* --force-eps */
    double force_kf;     /* --force-kf */
    double k0, energy;   /* --k0, --energy for -i K4 */
    uint64_t seed;       /* --seed */
    int device;          /* --device */
    int realspace;       /* write u and w real-space datasets (--no-realspace turns off) */
    int spectra;         /* write Spectra_HDF_Data.h5 (--no-spectra turns off) */
    int quirk;           /* --quirk-measures: reproduce the reference's measure arithmetic */
    double max_speed;    /* --max-speed blow-up guard */
    int transients;      /* --transients: the reference's -D__TRANSIENTS build (Solver/src/data_types.h:76-77):
                            the first TRANS_FRAC = 0.2 of the steps are not saved (Solver/src/solver.c:1339-1348) */
    int quiet;
} nsg_run;

int nsg_parse_args(int argc, char** argv, nsg_run* r);
void nsg_print_usage(const char* prog);

#endif
