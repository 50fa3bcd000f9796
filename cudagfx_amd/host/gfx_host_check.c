/*
 * gfx_host_check.c -- a CPU-only driver of the host services (input grammar,
 * ILDG reader/writer, Lehmer RNG, host observables) without the HIP library,
 * built with AddressSanitizer and UndefinedBehaviorSanitizer by
 * `make -C cudagfx_amd asan`.  It defines the lattice globals that
 * libcudagfx.so otherwise provides and runs one action per call:
 *
 *   gfx_host_check input <file>              read_input; prints the values
 *   gfx_host_check dims <file>               read_ildg_dims
 *   gfx_host_check read <file> lx ly lz lt   read_gf_ildg of an lx..lt field
 *   gfx_host_check write <file> lx ly lz lt prec seed
 *                                            random_init_gauge, write_gf_ildg
 *   gfx_host_check observe lx ly lz lt seed  random links and g, g_trafo,
 *                                            F, dAdA, plaquette
 *
 * Every action prints "rc=<status>" and exits 0 when it ran to the end, so a
 * sanitizer report (which aborts with a non-zero exit code) is the only way
 * for a run to fail.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/cudagfx_host.h"

int T = 0, L = 0, LX = 0, LY = 0, LZ = 0, VOLUME = 0;
double FUNC = 0.0, DADA = 0.0, maxDADA = 0.0, PLAQ = 0.0, BETA = 0.0;

static int set_dims(char **a) {
  LX = atoi(a[0]);
  LY = atoi(a[1]);
  LZ = atoi(a[2]);
  T = atoi(a[3]);
  L = LX;
  if (LX < 1 || LY < 1 || LZ < 1 || T < 1 || (long)LX * LY * LZ * T > (1L << 20)) return -1;
  VOLUME = LX * LY * LZ * T;
  return 0;
}

int main(int argc, char **argv) {
  if (argc < 3) {
    fprintf(stderr, "usage: %s input|dims|read|write|observe ...\n", argv[0]);
    return 2;
  }
  const char *act = argv[1];
  int rc = 0;
  if (!strcmp(act, "input")) {
    rc = read_input(argv[2]);
    printf("T=%d L=%d VOLUME=%d randseed=%d eps=%.9e maxit=%d ck=%d flags=%d%d%d\n", T, L, VOLUME,
           randseed, orxeps, orxmaxit, orxcheckinterval, orxflag, saflag, thermflag);
    printf("omega=%g precision=%d switch=%g gauge=%d mtseed=%ld micro=%d sampler=%d copies=%d files=%d\n",
           orxomega, orxprecision, orxswitch, orxgauge, mtseed, samicro, sasampler, orxcopies,
           norxfiles);
  } else if (!strcmp(act, "dims")) {
    int d[4] = {0, 0, 0, 0};
    rc = read_ildg_dims(argv[2], d);
    printf("dims=%d %d %d %d\n", d[0], d[1], d[2], d[3]);
  } else if (!strcmp(act, "read") && argc >= 7) {
    if (set_dims(argv + 3)) return 2;
    gfx_su3 *u = (gfx_su3 *)calloc((size_t)4 * VOLUME, sizeof(gfx_su3));
    if (!u) return 2;
    rc = read_gf_ildg(u, argv[2]);
    double s = 0.0;
    const double *p = (const double *)u;
    for (long k = 0; k < 72L * VOLUME; k++) s += p[k];
    printf("sum=%.17g\n", s);
    free(u);
  } else if (!strcmp(act, "write") && argc >= 9) {
    if (set_dims(argv + 3)) return 2;
    gfx_su3 *u = (gfx_su3 *)calloc((size_t)4 * VOLUME, sizeof(gfx_su3));
    if (!u) return 2;
    PlantSeeds(atol(argv[8]));
    random_init_gauge(u);
    PLAQ = mean_plaq(u);
    BETA = 6.0;
    rc = write_gf_ildg(u, argv[2], atoi(argv[7]));
    free(u);
  } else if (!strcmp(act, "observe") && argc >= 7) {
    if (set_dims(argv + 2)) return 2;
    gfx_su3 *u = (gfx_su3 *)calloc((size_t)4 * VOLUME, sizeof(gfx_su3));
    gfx_su3 *g = (gfx_su3 *)calloc((size_t)VOLUME, sizeof(gfx_su3));
    if (!u || !g) return 2;
    PlantSeeds(atol(argv[6]));
    random_init_gauge(u);
    random_init_trafo(g);
    g_trafo(u, g);
    const double F = gauge_functional(u), D = dAdA(u), P = mean_plaq(u);
    printf("F=%.17g dAdA=%.17g max=%.17g plaq=%.17g\n", F, D, maxDADA, P);
    free(u);
    free(g);
  } else {
    fprintf(stderr, "unknown action or missing arguments\n");
    return 2;
  }
  printf("rc=%d\n", rc);
  return 0;
}
