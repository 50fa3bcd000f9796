/*
 * cudagfx_host.h -- host services of libcudagfx.so that the reference's
 * driver (cudagaugefix.cu:599-828) calls around the device path: the input
 * grammar, ILDG gauge-file I/O, the Lehmer RNG and the host SU(3) routines.
 * Same names and argument meaning as the reference (file:line per entry).
 */
#ifndef CUDAGFX_HOST_H
#define CUDAGFX_HOST_H

#include "cudagfx.h"

#ifdef __cplusplus
extern "C" {
#endif

#ifndef GLOBAL_H
/* global.h:50-67 */
typedef struct SAPARAM_LIST {
  double Tmin;
  double Tmax;
  double expo;
  int N;
  int checkint;
} SAPARAM_LIST;
typedef struct THERMPARAM_LIST {
  int Nsweep;
  int Nsave;
  double beta;
  int checkint;
  int startcond;
} THERMPARAM_LIST;
#endif

/* ---- input grammar: read_input.l:29-46, 265-284 ---- */
extern int randseed, orxmaxit, orxcheckinterval, orxflag, saflag, thermflag;
extern double orxeps;
extern SAPARAM_LIST saparam;
extern THERMPARAM_LIST thermparam;
/* extension keys (accepted in addition to the reference grammar) */
extern double orxomega;       /* "omega = <float>"   in BeginORXInit (default 1.68) */
extern int orxprecision;      /* "precision = fp64 | mixed"                         */
extern double orxswitch;      /* "switch = <float>"  max dAdA to leave FP32         */
extern int orxgauge;          /* "gauge = landau | coulomb"                         */
extern long mtseed;           /* "MTSeed = <int>"    srand seed of the twisters     */
extern int samicro;           /* "micro = <int>"     in BeginSAInit                 */
extern int sasampler;         /* "sampler = reference | kp" in BeginSAInit          */
/* batch keys of BeginORXInit (no counterpart in the reference) */
#define GFX_MAX_GAUGEFILES 64
#define GFX_MAX_PATH 1024
extern int orxcopies;         /* "copies = <int>"    Gribov copies per configuration */
extern int norxfiles;         /* "gaugefiles = a b ..." configurations fixed with    */
extern char orxfiles[GFX_MAX_GAUGEFILES][GFX_MAX_PATH]; /* the -f file in one batch  */

/* read_input.l:265: returns 0, 2 if the file cannot be opened; a line the
 * grammar rejects prints "Parsing error in line %d" and returns 1 (the
 * reference exits with 1). Sets LX=LY=LZ=L, VOLUME. */
int read_input(char *conf_file);

/* ---- ILDG gauge files: gauge_io.c:245-595 ---- */
/* read 4V su3 links (32- or 64-bit, big-endian).  Returns 0, -1 on open or
 * record errors, 120 on a dimension mismatch with LX..T, 500 if the
 * precision entry does not match the data size (reference: exit(120|500)). */
int read_gf_ildg(gfx_su3 *gf, char *filename);
/* the lattice extents (lx, ly, lz, lt) of an ILDG file from its ildg-format
 * record, without reading the field (no counterpart in the reference: for
 * batches of configurations of different sizes).  0, or -1 if the file or
 * the record cannot be read. */
int read_ildg_dims(char *filename, int dims[4]);
/* write 4V links with prec = 32 or 64; 0 or -1 (open) / 500 (bad prec). */
int write_gf_ildg(gfx_su3 *gf, char *filename, const int prec);

/* ---- Lehmer RNG: rngs.c:51-150 ---- */
double Random(void);
void PlantSeeds(long x);
void PutSeed(long x);
void GetSeed(long *x);
void SelectStream(int index);

/* ---- host SU(3): su3manip.c:518-767 ---- */
void unit_init_trafo(gfx_su3 *trafofield);
void random_init_trafo(gfx_su3 *trafofield);
void unit_init_gauge(gfx_su3 *gaugefield);
void random_init_gauge(gfx_su3 *gaugefield);

/* ---- host observables: observables.c:17-242 ---- */
void g_trafo(gfx_su3 *u, gfx_su3 *g);
double gauge_functional(gfx_su3 *gf);
double dAdA(gfx_su3 *gf); /* sets maxDADA */
double mean_plaq(gfx_su3 *gf);

#ifdef __cplusplus
}
#endif
#endif
