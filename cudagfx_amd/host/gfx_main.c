/*
 * gfx_main.c -- the cudagaugefix driver (cudagaugefix.cu:599-828) on top of
 * libcudagfx.so: same command line (-i input -f gaugefile -d device), same
 * input grammar, same log lines, same output file "landau_<gaugefile>" and
 * the same exit codes (1 read error, -1 output exists, 300 no convergence,
 * 400 write error, 100/200 geometry/allocation).
 *
 * The extension keys of the input grammar select omega, the precision mode
 * (fp64 / mixed) and the gauge (landau / coulomb; the Coulomb output is
 * written as "coulomb_<gaugefile>").  With a BeginTHERMInit block the
 * driver thermalizes from a cold or hot start and writes the configuration
 * to the -f file name (the reference reads a file first and writes nothing,
 * heatbath_thermalization.cu / cudagaugefix.cu:687-705).
 */
#include <getopt.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>

#include <hip/hip_runtime_api.h>

#include "../../include/cudagfx_host.h"

extern int gfx_set_sa(int sampler, int nmicro, int ndir);

static void intro(void) {
  fprintf(stdout, "\n");
  fprintf(stdout, "########       cudagfx for AMD Instinct MI355X (gfx950)       ########\n");
  fprintf(stdout, "########       lattice Landau / Coulomb gauge fixing           ########\n\n\n");
}

static void usage(void) {
  fprintf(stdout, "Code to compute Landau gauge on gauge field\n");
  fprintf(stdout, "Usage:   cudagaugefix -i [inputfile] -f [gaugefile]\n");
  exit(0);
}

static void set_dims(const int d[4]) {
  LX = d[0];
  LY = d[1];
  LZ = d[2];
  T = d[3];
  L = LX;
  VOLUME = LX * LY * LZ * T;
}

/* The batch keys of BeginORXInit: every configuration (the -f file and the
 * `gaugefiles` list, each with its own extents from its ILDG header) is
 * fixed with `copies` starts in one batch -- the first start from
 * random_init_trafo as in the single run, the others Haar-random from the
 * twisters -- after simulated annealing of every start when a BeginSAInit
 * block is given.  For each configuration the copy with the largest F is
 * written as "landau_<file>".  Exit codes as the single run (1, -1, 300,
 * 400, 800 without the twister table). */
static int run_batch(const char *first) {
  const char *files[GFX_MAX_GAUGEFILES + 1];
  char outname[GFX_MAX_GAUGEFILES + 1][GFX_MAX_PATH + 16];
  int dims[GFX_MAX_GAUGEFILES + 1][4];
  int base[GFX_MAX_GAUGEFILES + 1];
  int nf = 0, rc;
  if (first[0]) files[nf++] = first;
  for (int k = 0; k < norxfiles; k++) files[nf++] = orxfiles[k];
  if (nf == 0) {
    printf("No gauge file given. Aborting...\n");
    return 1;
  }
  if (orxgauge != 0) {
    printf("copies and gaugefiles are implemented for gauge = landau. Aborting...\n");
    return 1;
  }
  const int ncopies = orxcopies > 1 ? orxcopies : 1;
  for (int f = 0; f < nf; f++) {
    struct stat st;
    if (read_ildg_dims((char *)files[f], dims[f]) != 0) {
      printf("Error reading configuration from %s\n", files[f]);
      return 1;
    }
    snprintf(outname[f], sizeof outname[f], "landau_%s", files[f]);
    if (stat(outname[f], &st) == 0) {
      printf(" output file %s is present, exiting...\n", outname[f]);
      return -1;
    }
  }
  printf("Batch of %d configuration(s) x %d start(s)\n", nf, ncopies);
  if ((rc = gfx_batch_create())) return rc;
  for (int f = 0; f < nf; f++) {
    set_dims(dims[f]);
    printf("%s: LX = %d, LY = %d, LZ = %d, T = %d\n", files[f], LX, LY, LZ, T);
    gfx_su3 *gf = (gfx_su3 *)malloc(sizeof(gfx_su3) * 4 * (size_t)VOLUME);
    gfx_su3 *tr = (gfx_su3 *)malloc(sizeof(gfx_su3) * (size_t)VOLUME);
    if (!gf || !tr) {
      fprintf(stderr, "host allocation failed\n");
      return 200;
    }
    if ((rc = read_gf_ildg(gf, (char *)files[f])) != 0) {
      printf("Error reading configuration from %s\n", files[f]);
      return rc == -1 ? 1 : rc;
    }
    random_init_trafo(tr);
    base[f] = gfx_batch_add(LX, LY, LZ, T, gf, tr);
    free(gf);
    free(tr);
    if (base[f] < 0) return -base[f];
    for (int c = 1; c < ncopies; c++) {
      const int q = gfx_batch_add_random_copy(base[f]);
      if (q < 0) {
        printf("Random Gribov copies need the twister table (%s). Aborting...\n", gfx_error_string(-q));
        return -q;
      }
    }
  }
  if (saflag == 1) {
    printf("Starting simulated annealing of every start...\n");
    printf("Tmin = %f, Tmax = %f, N = %d, expo = %f\n", saparam.Tmin, saparam.Tmax, saparam.N,
           saparam.expo);
    /* a mixed-precision batch anneals on FP32 links and g as well */
    gfx_batch_set_anneal_precision(orxprecision);
    if ((rc = gfx_batch_anneal(saparam.N, saparam.Tmin, saparam.Tmax, saparam.expo, samicro,
                               sasampler, 8)))
      return rc;
  }
  printf("Starting overrelaxation...\n");
  rc = gfx_batch_run(orxmaxit, orxeps, orxcheckinterval, orxprecision, orxswitch, 8);
  if (rc != GFX_OK && rc != GFX_ERR_NOCONV) return rc;
  int failed = 0;
  for (int f = 0; f < nf; f++) {
    for (int c = 0; c < ncopies; c++) {
      double F, D, M;
      int it;
      gfx_batch_result(base[f] + c, &F, &D, &M, &it);
      printf("%s copy %d: iter = %d\t F = %.16e \t dAdA = %.16e\t max(dAdA) = %.16e\n", files[f],
             c, it, F, D, M);
    }
    int qb;
    double Fb, D, M;
    int it;
    gfx_batch_best_copy(base[f], &qb, &Fb);
    gfx_batch_result(qb, &Fb, &D, &M, &it);
    if (it < 0) {
      printf("%s: gauge condition not reached.\n", files[f]);
      failed = 1;
      continue;
    }
    printf("%s: best copy %d, F = %.16e\n", files[f], qb - base[f], Fb);
    set_dims(dims[f]);
    gfx_su3 *gf = (gfx_su3 *)malloc(sizeof(gfx_su3) * 4 * (size_t)VOLUME);
    gfx_su3 *tr = (gfx_su3 *)malloc(sizeof(gfx_su3) * (size_t)VOLUME);
    if (!gf || !tr) {
      fprintf(stderr, "host allocation failed\n");
      return 200;
    }
    if (read_gf_ildg(gf, (char *)files[f]) != 0 || gfx_batch_get_trafo(qb, tr) != GFX_OK) {
      printf("Error reading configuration from %s\n", files[f]);
      return 1;
    }
    g_trafo(gf, tr);
    PLAQ = mean_plaq(gf);
    DADA = dAdA(gf);
    FUNC = gauge_functional(gf);
    printf("Final HOST values:\n");
    printf("PLAQ = %.16f\n", PLAQ);
    printf("F = %.16e \t dAdA = %.16e\t max(dAdA) = %.16e\n", FUNC, DADA, maxDADA);
    printf("Writing out the gauge fixed field ...");
    rc = write_gf_ildg(gf, outname[f], 64);
    free(gf);
    free(tr);
    if (rc != 0) {
      fprintf(stderr, "Error writing gauge field. Aborting...\n");
      return 400;
    }
    printf("done.\n");
  }
  gfx_batch_destroy();
  finalize_MT();
  return failed ? 300 : 0;
}

int main(int argc, char *argv[]) {
  char inputfilename[1024] = "input.in";
  char gaugefilename[1024] = "";
  char fixedgaugename[1100] = "";
  int dev = 0, c, ret;
  intro();
  while ((c = getopt(argc, argv, "h?:i:d:f:")) != -1) {
    switch (c) {
      case 'i':
        snprintf(inputfilename, sizeof inputfilename, "%s", optarg);
        printf("The input file is: %s\n", inputfilename);
        break;
      case 'd':
        dev = atoi(optarg);
        break;
      case 'f':
        snprintf(gaugefilename, sizeof gaugefilename, "%s", optarg);
        printf("The gauge  file is: %s\n", gaugefilename);
        break;
      case 'h':
      case '?':
      default:
        usage();
        break;
    }
  }
  printf("setting device to %d\n", dev);
  fflush(stdout);
  if (hipSetDevice(dev) != hipSuccess) {
    fprintf(stderr, "cannot select device %d\n", dev);
    exit(700);
  }

  ret = read_input(inputfilename);
  if (ret == 2) {
    fprintf(stderr, "cannot open input file %s\n", inputfilename);
    exit(1);
  }
  if (ret != 0) exit(1);
  printf("LX = %d, LY = %d, LZ = %d, T = %d\n", LX, LY, LZ, T);
  snprintf(fixedgaugename, sizeof fixedgaugename, "%s%s", orxgauge ? "coulomb_" : "landau_",
           gaugefilename);
  printf("The fixed gauge  file is: %s\n", fixedgaugename);
  gfx_set_omega(orxomega);
  gfx_set_precision(orxprecision, orxswitch);
  gfx_set_mt_seed(mtseed);
  gfx_set_sa(sasampler, samicro, orxgauge ? 3 : 4);

  gfx_su3 *g_gf = (gfx_su3 *)malloc(sizeof(gfx_su3) * 4 * (size_t)VOLUME);
  gfx_su3 *trafo1 = (gfx_su3 *)malloc(sizeof(gfx_su3) * (size_t)VOLUME);
  if (!g_gf || !trafo1) {
    fprintf(stderr, "host allocation failed\n");
    exit(200);
  }

  printf("Setting random seed to %d\n", randseed);
  PlantSeeds(randseed);

  if (thermflag == 1) {
    if (thermparam.startcond == 0)
      unit_init_gauge(g_gf);
    else
      random_init_gauge(g_gf);
    if ((ret = init_thermalization(g_gf))) exit(ret);
    calc_plaquette(1);
    thermalize_gauge();
    get_gauge(g_gf);
    PLAQ = calc_plaquette(1);
    BETA = thermparam.beta;
    finalize_MT();
    finalize_thermalization();
    if (gaugefilename[0]) {
      struct stat st;
      if (stat(gaugefilename, &st) == 0) {
        printf(" output file %s is present, exiting...\n", gaugefilename);
        exit(-1);
      }
      if (write_gf_ildg(g_gf, gaugefilename, 64) != 0) {
        fprintf(stderr, "Error writing gauge field. Aborting...\n");
        exit(400);
      }
    }
    free(g_gf);
    free(trafo1);
    return 0;
  }

  if (orxflag == 1 && (orxcopies > 1 || norxfiles > 0)) {
    free(g_gf);
    free(trafo1);
    ret = run_batch(gaugefilename);
    exit(ret);
  }

  if (read_gf_ildg(g_gf, gaugefilename) != 0) {
    printf("Error reading configuration from %s\n", gaugefilename);
    exit(1);
  }
  {
    struct stat st;
    if (stat(fixedgaugename, &st) == 0) {
      printf(" output file %s is present, exiting...\n", fixedgaugename);
      exit(-1);
    }
  }

  random_init_trafo(trafo1);
  if ((ret = init_gaugefixing(g_gf, trafo1))) exit(ret);

  calc_plaquette(1);
  {
    const double F = gauge_functional(g_gf);
    const double dada = dAdA(g_gf);
    printf("HOST FUNC = %.16e\tHOST dAdA = %.16e\n", F, dada);
  }

  if (saflag == 1) {
    printf("Starting simulated annealing...\n");
    printf("Tmin = %f, Tmax = %f, N = %d, expo = %f\n", saparam.Tmin, saparam.Tmax, saparam.N,
           saparam.expo);
    simannealing_gauge();
  }

  if (orxflag == 1) {
    printf("Starting overrelaxation...\n");
    if (orxgauge == 0) {
      ret = overrelax_gauge(orxmaxit, orxeps, orxcheckinterval);
    } else {
      int *iters = (int *)malloc(sizeof(int) * T);
      ret = gfx_coulomb_gauge(orxmaxit, orxeps, orxcheckinterval, iters) == GFX_OK ? 0 : -1;
      free(iters);
    }
    if (ret < 0) {
      printf("Gauge condition not reached. Aborting...\n");
      finalize_gaugefixing();
      free(trafo1);
      free(g_gf);
      exit(300);
    }
  }
  calc_plaquette(1);

  printf("Transferring back to host...\n");
  printf("Applying trafo on host...\n");
  get_trafo(trafo1);
  apply_trafo(g_gf);
  PLAQ = mean_plaq(g_gf);
  DADA = dAdA(g_gf);
  FUNC = gauge_functional(g_gf);
  printf("Final HOST values:\n");
  printf("PLAQ = %.16f\n", PLAQ);
  printf("F = %.16e \t dAdA = %.16e\t max(dAdA) = %.16e\n", FUNC, DADA, maxDADA);

  printf("Writing out the gauge fixed field ...");
  ret = write_gf_ildg(g_gf, fixedgaugename, 64);
  if (ret != 0) {
    fprintf(stderr, "Error writing gauge field. Aborting...\n");
    exit(400);
  }
  printf("done.\n");
  finalize_MT();
  finalize_gaugefixing();
  free(trafo1);
  free(g_gf);
  return 0;
}
