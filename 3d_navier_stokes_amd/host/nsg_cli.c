/*
 * nsg_cli.c -- command line of the host driver.  The short options, their meaning, defaults and
 * validation errors follow GetCMLArgs (Solver/src/utils.c:34-237): the same getopt string
 * "o:h:n:d:s:e:t:v:i:c:p:f:z:" and the same checks (N > 2 and even, t0 >= 0, T >= t0, dt > 0,
 * CFL > 0, nu > 0).  Compile-time switches of the reference (Solver/Makefile:46-55,
 * Solver/src/data_types.h:51-115) are run-time long options here.
 */
#define _GNU_SOURCE
#include <getopt.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "nsg_run.h"

void nsg_print_usage(const char* prog) {
    fprintf(stderr,
            "usage: %s [reference flags] [--long options]\n"
            "  -o DIR     output directory (give -o twice for file-only output)\n"
            "  -n N       grid size per dimension (give three times: -n 64 -n 64 -n 64)\n"
            "  -s T0      start time            -e T    end time         -h DT  time step\n"
            "  -c CFL     CFL constant          -v NU   viscosity        -p K   save every K steps\n"
            "  -i IC      TAYLOR_GREEN | SHAPIRO | RANDOM | TESTING | K4 (other values: random)\n"
            "  -t TAG     output directory tag  -f F    NONE | ZERO | CONST_POWER\n"
            "  -z FILE    restart from the last /Iter_* group of a Main_HDF_Data.h5 file\n"
            "  -d ALPHA   Ekman drag (accepted and unused, as in the reference)\n"
            "  --visc-pow P   2: nu|k|^4 (default, as the reference's `make solver`, -D__HYPER)\n"
            "                 1: nu|k|^2 (the reference's test build, no -D__HYPER)\n"
            "  --transients   do not save the first 20%% of the steps (reference -D__TRANSIENTS)\n"
            "  --scheme S     rk4cn (reference) | ifrk4 | rk5cn (-D__RK5) | dp5 (-D__DPRK5, adaptive)\n"
            "  --precision P  fp64 | fp32bulk\n"
            "  --adaptive     CFL-adaptive step with the -c constant\n"
            "  --force-eps E  constant-power forcing rate (with -f CONST_POWER; default 0.1)\n"
            "  --force-kf K   forcing band |k| <= K (default 2.5)\n"
            "  --k0 K --energy E --seed S   parameters of -i K4\n"
            "  --device D  --max-speed U  --no-realspace  --no-spectra  --quirk-measures  --quiet\n",
            prog);
}

static int fail(const char* msg, double v) {
    fprintf(stderr, "\n[ERROR] Parsing of Command Line Arguments Failed: %s [%g]\n-->> Exiting!\n\n", msg, v);
    return 1;
}

int nsg_parse_args(int argc, char** argv, nsg_run* r) {
    memset(r, 0, sizeof *r);
    strncpy(r->output_dir, "./Data/Tmp/", sizeof r->output_dir - 1);
    strncpy(r->output_tag, "NO_TAG", sizeof r->output_tag - 1);
    strncpy(r->input_file, "NONE", sizeof r->input_file - 1);
    strncpy(r->u0, "TG_VORT", sizeof r->u0 - 1);
    strncpy(r->forcing, "NONE", sizeof r->forcing - 1);
    r->N[0] = r->N[1] = r->N[2] = 64;
    r->t0 = 0.0;
    r->dt = 1e-3;
    r->T = 1.0;
    r->cfl = sqrt(3.0);
    r->nu = 1.0;
    r->save_every = 100;
    /* the reference's production build `make solver` is hyperviscous (SOLVER_FLAGS -D__HYPER,
     * Solver/Makefile:46); its test build (TEST_FLAGS, Solver/Makefile:55) is not: --visc-pow 1 */
    r->visc_pow = 2;
    r->scheme = NSG_SCHEME_RK4CN;
    r->precision = NSG_FP64;
    r->force_eps = 0.1;
    r->force_kf = 2.5;
    r->k0 = 4.0;
    r->energy = 0.5;
    r->seed = 1;
    r->realspace = 1;
    r->spectra = 1;

    static struct option longopts[] = {
        {"visc-pow", required_argument, 0, 1000}, {"scheme", required_argument, 0, 1001},
        {"precision", required_argument, 0, 1002}, {"adaptive", no_argument, 0, 1003},
        {"force-eps", required_argument, 0, 1004}, {"force-kf", required_argument, 0, 1005},
        {"k0", required_argument, 0, 1006}, {"energy", required_argument, 0, 1007},
        {"seed", required_argument, 0, 1008}, {"device", required_argument, 0, 1009},
        {"no-realspace", no_argument, 0, 1010}, {"no-spectra", no_argument, 0, 1011},
        {"quirk-measures", no_argument, 0, 1012}, {"max-speed", required_argument, 0, 1013},
        {"quiet", no_argument, 0, 1014}, {"help", no_argument, 0, 1015},
        {"transients", no_argument, 0, 1016}, {0, 0, 0, 0}};
    int dim = 0, outdir_seen = 0, c;
    optind = 1;
    while ((c = getopt_long(argc, argv, "o:h:n:d:s:e:t:v:i:c:p:f:z:", longopts, NULL)) != -1) {
        switch (c) {
            case 'o':
                if (outdir_seen == 0) strncpy(r->output_dir, optarg, sizeof r->output_dir - 1);
                else r->file_only = 1;
                outdir_seen++;
                break;
            case 'n': {
                long n = atol(optarg);
                if (dim > 2) return fail("too many -n values", (double)n);
                if (n <= 2) return fail("The minimum dimension size must be greater than 2", (double)n);
                if (n % 2 != 0) return fail("The system dimension must be a multiple of 2", (double)n);
                r->N[dim++] = (int)n;
                break;
            }
            case 's':
                r->t0 = atof(optarg);
                if (r->t0 < 0) return fail("The integration start time must be positive", r->t0);
                break;
            case 'e':
                r->T = atof(optarg);
                if (r->T < 0) return fail("The integration end time must be positive", r->T);
                if (r->T < r->t0) return fail("The end time must be greater than the initial time", r->T);
                break;
            case 'h':
                r->dt = atof(optarg);
                if (r->dt <= 0) return fail("The timestep must be strictly positive", r->dt);
                break;
            case 'c':
                r->cfl = atof(optarg);
                if (r->cfl <= 0) return fail("The CFL constant must be strictly positive", r->cfl);
                break;
            case 'v':
                r->nu = atof(optarg);
                if (r->nu <= 0) return fail("The viscosity must be strictly positive", r->nu);
                break;
            case 'd':
                r->ekman = atof(optarg);
                if (r->ekman <= 0) return fail("The Ekman friction must be strictly positive", r->ekman);
                break;
            case 'i':
                if (!strcmp(optarg, "TAYLOR_GREEN") || !strcmp(optarg, "SHAPIRO") || !strcmp(optarg, "TESTING") ||
                    !strcmp(optarg, "RANDOM") || !strcmp(optarg, "K4"))
                    strncpy(r->u0, optarg, sizeof r->u0 - 1);
                else
                    strncpy(r->u0, "NONE", sizeof r->u0 - 1);  /* -> random, as in the reference */
                break;
            case 't': strncpy(r->output_tag, optarg, sizeof r->output_tag - 1); break;
            case 'p': r->save_every = atoi(optarg); break;
            case 'f':
                if (!strcmp(optarg, "ZERO") || !strcmp(optarg, "CONST_POWER")) strncpy(r->forcing, optarg, sizeof r->forcing - 1);
                else strncpy(r->forcing, "NONE", sizeof r->forcing - 1);
                break;
            case 'z':
                strncpy(r->input_file, optarg, sizeof r->input_file - 1);
                if (access(r->input_file, F_OK) != 0) {
                    fprintf(stderr, "\n[ERROR] The input file [%s] cannot be found\n", r->input_file);
                    return 1;
                }
                break;
            case 1000: r->visc_pow = atoi(optarg); break;
            case 1001:
                if (!strcmp(optarg, "rk4cn")) r->scheme = NSG_SCHEME_RK4CN;
                else if (!strcmp(optarg, "ifrk4")) r->scheme = NSG_SCHEME_IFRK4;
                else if (!strcmp(optarg, "rk5cn")) r->scheme = NSG_SCHEME_RK5CN;
                else if (!strcmp(optarg, "dp5")) r->scheme = NSG_SCHEME_DP5;
                else return fail("unknown --scheme", 0);
                break;
            case 1002:
                if (!strcmp(optarg, "fp64")) r->precision = NSG_FP64;
                else if (!strcmp(optarg, "fp32bulk")) r->precision = NSG_FP32BULK;
                else return fail("unknown --precision", 0);
                break;
            case 1003: r->adaptive = 1; break;
            case 1004: r->force_eps = atof(optarg); break;
            case 1005: r->force_kf = atof(optarg); break;
            case 1006: r->k0 = atof(optarg); break;
            case 1007: r->energy = atof(optarg); break;
            case 1008: r->seed = strtoull(optarg, NULL, 10); break;
            case 1009: r->device = atoi(optarg); break;
            case 1010: r->realspace = 0; break;
            case 1011: r->spectra = 0; break;
            case 1012: r->quirk = 1; break;
            case 1013: r->max_speed = atof(optarg); break;
            case 1014: r->quiet = 1; break;
            case 1015: nsg_print_usage(argv[0]); return 2;
            case 1016: r->transients = 1; break;
            default:
                fprintf(stderr, "\n[ERROR] Incorrect command line flag encountered\n");
                nsg_print_usage(argv[0]);
                return 1;
        }
    }
    if (r->save_every <= 0) return fail("-p must be positive", r->save_every);
    if (r->visc_pow < 1 || r->visc_pow > 4) return fail("--visc-pow must be 1..4", r->visc_pow);
    return 0;
}
