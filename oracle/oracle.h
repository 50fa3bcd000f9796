/*
 * oracle.h -- CPU restatement of the cudagfx gauge-fixing path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links this code;
 * it is loaded by tests/, by __graft_entry__.smoke() to check the GPU result and by
 * bench.py as the CPU baseline.  Every function restates the arithmetic of
 * the reference (files cited per function, paths relative to the reference
 * checkout) in the same operation order, on the same 2-row (12-real) storage
 * the reference keeps on the device, so that a GPU result can be compared
 * element by element.
 *
 * Storage conventions (identical to the reference's host/device contract):
 *   su3 host matrix      : 18 doubles, row-major, (re,im) pairs
 *                          (global.h:26-33)
 *   2-row matrix         : 12 doubles = rows 0 and 1 of the su3 matrix
 *                          (su3manip.c:212-230, cudaglobal.h:25-28)
 *   lexicographic site s : x + LX*(y + LY*(z + LZ*t)), mu = 0..3 = x,y,z,t
 *                          (cudagaugefix.cu:79-85)
 *   gauge link l         : 4*s + mu (observables.c:34)
 */
#ifndef CUDAGFX_ORACLE_H
#define CUDAGFX_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- geometry ---------------- */
/* cudagaugefix.cu:72-118 (initnn): nn[8*s+mu] forward, nn[8*s+4+mu] backward */
void orc_init_nn(int LX, int LY, int LZ, int LT, int *nn);
/* cudagaugefix.cu:129-196 (initnn_eo): lexicographic order filtered by parity */
void orc_init_eo(int LX, int LY, int LZ, int LT, int *eo_even, int *eo_odd);

/* ---------------- Lehmer RNG (rngs.c) ---------------- */
double orc_lehmer_random(void);          /* rngs.c:51-68  */
void   orc_lehmer_plant_seeds(long x);   /* rngs.c:71-98  */
void   orc_lehmer_select_stream(int i);  /* rngs.c:140-150 */
long   orc_lehmer_get_seed(void);        /* rngs.c:129-137 */
void   orc_lehmer_put_seed(long x);      /* rngs.c:101-126 (x>0 only) */

/* su3manip.c:539-641: ten Lehmer draws per site, Gram-Schmidt, row 2 = conj(r0 x r1) */
void orc_random_init_trafo(double *g_su3, int V);
/* su3manip.c:665-767: same construction for 4V links */
void orc_random_init_gauge(double *u_su3, int V);

/* ---------------- layout conversions ---------------- */
void orc_su3_to_2v(const double *su3, double *two_v, long n);      /* su3manip.c:212-230 */
void orc_2v_to_su3(const double *two_v, double *su3, long n);      /* su3manip.c:303-359 */

/* ---------------- the sweep (overrelax.cu) ---------------- */
/*
 * One red-black half sweep, in place on g2v (12 doubles per site):
 * overrelax.cu:335-446 (dev_overrelax_step).  `sites` lists the sites of the
 * parity being updated (eo_even or eo_odd), `nsites` = V/2.  ndir = 4 for
 * Landau, 3 for Coulomb (spatial directions only).  `omega` is the
 * overrelaxation parameter (1.68 in the reference, overrelax.cu:416) and
 * `nterms` the truncation N of the binomial series (3 in the reference).
 * `active` (may be NULL) marks time slices that still update (Coulomb).
 * Uses OpenMP over the sites when compiled with it.
 */
void orc_or_halfsweep(const double *u2v, double *g2v, const int *nn,
                      const int *sites, int nsites, int ndir,
                      double omega, int nterms,
                      const unsigned char *active, int slice_volume);

/*
 * overrelax.cu:482-605 + 610-671 (dev_functional + calc_functional).
 * Returns maxdAdA; writes F = sum Re tr g U g^+ / (ndir*V) and
 * dAdA = sum Re tr Delta Delta^+ / V.  Block partial sums of 128 sites are
 * formed first, then summed in block order, as the reference does.
 * ndir = 4 (Landau) or 3 (Coulomb).
 */
double orc_functional(const double *u2v, const double *g2v, const int *nn,
                      int V, int ndir, double *F, double *dAdA);

/* Per time slice variant for Coulomb gauge: arrays of length LT. */
void orc_functional_slices(const double *u2v, const double *g2v, const int *nn,
                           int V, int LT, int ndir,
                           double *F, double *dAdA, double *maxdAdA);

/*
 * overrelax.cu:677-823 (overrelax_gauge): loop of even+odd half sweeps,
 * check every `ck` iterations, stop when maxdAdA < eps.  Returns the stop
 * iteration i, or -1.  hist (may be NULL) receives up to hist_cap triples
 * (F, dAdA, maxdAdA) in check order; *nhist the count.
 */
int orc_overrelax_gauge(const double *u2v, double *g2v, int LX, int LY, int LZ,
                        int LT, int maxit, double eps, int ck, double omega,
                        double *hist, int hist_cap, int *nhist);

/* Coulomb gauge: per time slice stop; iters[t] = stop iteration or -1. */
int orc_coulomb_gauge(const double *u2v, double *g2v, int LX, int LY, int LZ,
                      int LT, int maxit, double eps, int ck, double omega,
                      int *iters);

/* overrelax.cu:34-111 (dev_mean_plaq) on 2-row links; reduction over t as
 * calc_plaquette (overrelax.cu:115-138). */
double orc_plaquette_2v(const double *u2v, const int *nn, int LX, int LY,
                        int LZ, int LT);

/* ---------------- host observables (observables.c) ---------------- */
void   orc_g_trafo(double *u_su3, const double *g_su3, const int *nn, int V); /* 17-47 */
double orc_gauge_functional(const double *u_su3, int V);                     /* 53-62 */
double orc_dAdA(const double *u_su3, const int *nn, int V, double *maxdAdA); /* 70-140 */
double orc_mean_plaq(const double *u_su3, const int *nn, int V);             /* 176-242 */

/* ---------------- small SU(3) helpers exposed for tests ---------------- */
void orc_su3_normalize(double *m18);                  /* dev_su3.cu:1569-1634 */
void orc_cm_relax(double *g18, const double *star18); /* overrelax.cu:203-269 */
void orc_overrelax_step(double *g18, const double *gn18, double w, int N);
/* the short form of the 3-term series (the library's polish sweeps), then
 * the normalisation */
void orc_overrelax_step_short(double *g18, const double *h18, double w); /* 149-177 */
void orc_reconstruct(const double *r12, double *m18);         /* dev_su3.cu:737-820 */
void orc_reconstruct_dagger(const double *r12, double *m18);  /* dev_su3.cu:826-929 */

/* ---------------- Mersenne Twister (MersenneTwister.cu) ---------------- */
typedef struct {
  uint32_t matrix_a, mask_b, mask_c, seed;
} orc_mt_params;

/*
 * MersenneTwister.cu:142-199 (RandomGPU) for twisters [0, count): writes
 * out[i + k*count] for k < nper as ((float)x + 1.0f)/2^32, keeps the
 * 19-word states in `state` (count x 19) and params[i].seed = state[0]
 * after the call.  initialized = 0 seeds the state from params[i].seed with
 * the 1812433253 recurrence; 1 reloads `state` and restarts at index 0
 * (the reference's restart semantics).  out_words (may be NULL) receives
 * the tempered 32-bit words in the same layout.
 */
void orc_mt_generate(orc_mt_params *params, uint32_t *state, int count,
                     int nper, int initialized, float *out, uint32_t *out_words);

/* Generic MT engine for known-answer checks: MT19937 (n=624, m=397, r=31). */
uint32_t orc_mt19937_nth(uint32_t seed, int n);

/* MersenneTwister.cu:210-229 restated in double precision: the exact
 * Box-Muller pair for (u1, u2), used to bound the GPU float results.  The
 * reference's float path (logf, __cosf, __sinf) is not reproduced bit for
 * bit: parity unpinned at the last bit, bounded to 2e-5 instead. */
void orc_box_muller_exact(double u1, double u2, double *z1, double *z2);

/* iAlignUp(iDivUp(n, count), 2) (MersenneTwister.cu:246,252) */
int orc_mt_nper(int n, int count);

#ifdef __cplusplus
}
#endif
#endif
