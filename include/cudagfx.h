/*
 * cudagfx.h -- C ABI of libcudagfx.so, the MI355X (gfx950) implementation of
 * the cudagfx gauge-fixing path.
 *
 * The first block carries the reference's own entry points, same names,
 * same argument meaning (reference file:line cited per declaration); a
 * driver written against the reference's cudagaugefix.cu calls them
 * unchanged.  Differences are stated per function: entry points that the
 * reference declared `void` and ended with exit(code) return that code as
 * an int instead, and functions that took device pointers take none (the
 * library owns the device fields between init_* and finalize_*).
 *
 * Host matrices use the reference layout: su3 = 3x3 complex double,
 * row-major (global.h:26-33); gauge field gf[4*s+mu], trafo g[s] with
 * s = x + LX*(y + LY*(z + LZ*t)) (cudagaugefix.cu:79-85).
 *
 * Single instance, global state, one host thread, as the reference.  Every
 * value is final on return.
 */
#ifndef CUDAGFX_H
#define CUDAGFX_H

#include <stddef.h>
#include <stdint.h>

#ifdef GLOBAL_H
/* the reference's global.h came first: use its su3 type */
typedef su3 gfx_su3;
#else
typedef struct {
  double re, im;
} gfx_complex;
typedef gfx_complex gfx_su3[3][3];
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes: the reference's exit codes (cudagaugefix.cu, overrelax.cu) ---- */
#define GFX_OK 0
#define GFX_ERR_GEOMETRY 100    /* overrelax.cu:687-690: (V/2) % BLOCK, shape rules   */
#define GFX_ERR_ALLOC 200       /* cudagaugefix.cu:216-219                            */
#define GFX_ERR_NOCONV 300      /* cudagaugefix.cu:741-748                            */
#define GFX_ERR_WRITE 400       /* cudagaugefix.cu:808-811                            */
#define GFX_ERR_STATE 600       /* called before init_* or after finalize_*           */
#define GFX_ERR_DEVICE 700      /* a HIP runtime call failed                          */
#define GFX_ERR_MTDAT 800       /* MersenneTwister.cu:81-90 (.dat missing or short)   */
#define GFX_ERR_SIZE 900        /* input larger than a fixed-size device array allows */

/* ---- lattice globals (global.h:77, 86-94), defined by the library ---- */
extern int T, L, LX, LY, LZ, VOLUME;
extern double FUNC, DADA, maxDADA, PLAQ, BETA;

/* ================= reference entry points ================= */

/* cudagaugefix.cu:204 -- copy U (4V su3) and g (V su3) to the device in the
 * even/odd structure-of-arrays layout.  Returns GFX_OK, GFX_ERR_GEOMETRY or
 * GFX_ERR_ALLOC (reference: void + exit(100|200)). */
int init_gaugefixing(gfx_su3 *gf, gfx_su3 *trafo);

/* cudagaugefix.cu:381 */
void finalize_gaugefixing(void);

/* cudagaugefix.cu:420 -- device copy of U for thermalization. */
int init_thermalization(gfx_su3 *gf);

/* cudagaugefix.cu:556 */
void finalize_thermalization(void);

/* overrelax.cu:677 -- red-black overrelaxation until max_x dAdA(x) < eps,
 * checked every `checkinterval` iterations; returns the stop iteration or
 * -1.  Prints the reference's "iter" / "CONVERGENCE!" lines
 * (overrelax.cu:782, 795-798) unless gfx_set_verbose(0).  Uses the
 * precision mode of gfx_set_precision and omega of gfx_set_omega. */
int overrelax_gauge(int maxit, double eps, int checkinterval);

/* overrelax.cu:610 -- F, dAdA, max dAdA of the current g; sets FUNC, DADA,
 * maxDADA and returns maxDADA.  (reference: took the device fields as
 * arguments; here the library's own fields are used.) */
double calc_functional(void);

/* overrelax.cu:115 -- mean plaquette of the device links; prints
 * "PLAQ = %.16f" when outputyn == 1. */
double calc_plaquette(int outputyn);

/* simulated_annealing.cu:309 -- SA on g with the schedule in saparam. */
void simannealing_gauge(void);

/* heatbath_thermalization.cu:568 -- link heatbath with thermparam. */
void thermalize_gauge(void);

/* MersenneTwister.cu:236 -- load MersenneTwister.dat from the working
 * directory, seed the 4096 twisters from srand(seed)/rand() (seed:
 * gfx_set_mt_seed, else time(NULL) as the reference) and fill the gauss and
 * uniform fields.  Returns GFX_ERR_MTDAT if the file is missing. */
int init_MT(int n_gaussnumbers, int n_unifnumbers);
/* MersenneTwister.cu:295 */
void update_MT(void);
/* MersenneTwister.cu:327 */
void finalize_MT(void);

/* ================= replacements for direct device-pointer use ================= */

/* cudagaugefix.cu:789-792 (cudaMemcpy D2H of dev_trafo1 + from2vtosu3_trafo):
 * current g as V host su3 matrices, row 2 rebuilt from rows 0,1. */
int get_trafo(gfx_su3 *trafo);
/* set g from V host su3 matrices (rows 0,1 are used). */
int set_trafo(gfx_su3 *trafo);
/* current device links as 4V host su3 matrices (row 2 rebuilt). */
int get_gauge(gfx_su3 *gf);
/* observables.c:17-47 (g_trafo) on the device: gf[4s+mu] <- g(s) gf g^+(s+mu)
 * for host links gf (all three rows used), with the current g. */
int apply_trafo(gfx_su3 *gf);

/* ================= extensions (no counterpart in the reference) ================= */

#define GFX_PREC_FP64 0   /* plain FP64 path                                    */
#define GFX_PREC_MIXED 1  /* FP32 12-real bulk sweeps, FP64 reductions + polish */
/* not precision modes: the two sweep forms of a mixed run on packed
 * links (gfx_set_packed_links), as the `precision` argument of
 * gfx_time_sweeps / gfx_get_sweep_variant / gfx_set_sweep_variant:
 * PACKED64 the FP64 polish (FP64 g on 24-bit high + 16-bit low parts),
 * PACKED32 the FP32 phase (FP32 g on the 21-bit links) */
#define GFX_SWEEP_PACKED64 2
#define GFX_SWEEP_PACKED32 3
/* the FP64 polish of the mixed Landau run: links and g in 32-bit fixed
 * point (int32 = rint(u 2^31) per real, 48 bytes per matrix), FP64 arithmetic */
#define GFX_SWEEP_FIXED64 4

#define GFX_GAUGE_LANDAU 0
#define GFX_GAUGE_COULOMB 1

/* overrelaxation parameter; the reference hard-codes 1.68 (overrelax.cu:416) */
int gfx_set_omega(double omega);
/* precision mode; switch_theta = max dAdA at which the mixed path leaves
 * FP32 (default 1e-6) */
int gfx_set_precision(int mode, double switch_theta);
/* the mixed path also leaves FP32 when its best max dAdA is below `gate`
 * and has not halved for `iters` iterations (the FP32 rounding floor, near
 * 5e-11 on 32^4); iters = 0 turns this off.  Default 300, 1e-8. */
int gfx_set_fp32_stall(int iters, double gate);
/* links read by the sweeps of a mixed Landau, Coulomb or batch run: 1 (default)
 * packed links.  The FP32 phase reads each real u of a link as the 21-bit
 * integer rint(u 2^20) times 2^-20 (within 2^-21 of u, 32 instead of 48 bytes
 * per link).  The FP64 polish holds links
 * and g in 32-bit fixed point (rint(u 2^31) per real, 48 bytes per matrix,
 * FP64 arithmetic): Landau runs for eps > 1e-15, batch runs and Coulomb runs
 * whose slices hold whole workgroups for eps > 3.1e-13; otherwise it reads
 * a 24-bit hi = rint(u 2^23) plus an int16 lo = rint((u - hi 2^-23) 2^38),
 * i.e. hi 2^-23 + lo 2^-38 (within 2^-38 of u, 60 instead of 96 bytes per
 * link), with FP64 g.  A Landau stop found in FP32 or on packed or fixed-point
 * fields is confirmed by a functional on the FP64 links (the sweeps go on if
 * it does not hold); an FP32 check never stops a Coulomb slice or batch
 * lattice (it moves to FP64), and an FP64 one stops it only when the
 * rounding's worst effect on Delta keeps its max dAdA below eps (packed
 * links for eps > 3e-17 only).  A batch run's reported F, dAdA and max dAdA
 * are those of its returned g on the FP64 links.  0: the FP32 and FP64
 * links. */
int gfx_set_packed_links(int on);
/* 0 silences the per-check log lines */
int gfx_set_verbose(int verbose);
/* seed for srand() before the twister seeds are drawn (negative: time(NULL)) */
int gfx_set_mt_seed(long seed);

/* Coulomb gauge: every time slice is gauge fixed over the 3 spatial
 * directions; a slice stops updating once its own max dAdA < eps (checked
 * every checkinterval).  iters[t] receives the stop iteration of slice t
 * (-1 if not reached).  Returns GFX_OK or GFX_ERR_NOCONV. */
int gfx_coulomb_gauge(int maxit, double eps, int checkinterval, int *iters);
/* per-slice F, dAdA, max dAdA (arrays of length T) for the current g */
int gfx_functional_slices(double *F, double *dAdA, double *maxdAdA);

/* one red-black half sweep of the given parity (0 even, 1 odd) in the current
 * precision mode; ndir 4 (Landau) or 3 (Coulomb, all slices) */
int gfx_halfsweep(int parity, int ndir);

/* Timing of the plain sweep: runs `nsweeps` full Landau sweeps (even + odd)
 * on the current g in `precision` (GFX_PREC_FP64, GFX_PREC_MIXED for the FP32
 * sweep on FP32 links, GFX_SWEEP_PACKED32 / PACKED64 for the packed forms)
 * and returns the mean kernel
 * time of one half-sweep launch in ms, measured with HIP events on the
 * library's stream. */
int gfx_time_sweeps(int nsweeps, int precision, double *ms_per_halfsweep);
/* wall-clock of the last overrelax_gauge / gfx_coulomb_gauge call in ms
 * (first sweep to the stop decision), and the number of FP32 and FP64
 * iterations it ran. */
int gfx_last_run_stats(double *ms, int *iters_fp32, int *iters_fp64);
/* Launch times inside real runs: with `on`, each run of consecutive
 * half-sweep launches of one form without a fused check (the launches
 * between two checks of overrelax_gauge) is bracketed by a pair of HIP
 * events on the library's stream (at most 2048 such segments; the record
 * restarts with each overrelax_gauge call and with this call). */
int gfx_set_sweep_sampling(int on);
/* Mean time per launch in ms over the bracketed launches of one sweep form
 * (`precision` as for gfx_time_sweeps), the smallest segment mean, and how
 * many launches there were; waits for the last segment. */
int gfx_sweep_sample_stats(int precision, double *mean_ms, double *min_ms, int *count);
/* wait for all device work of the library */
int gfx_sync(void);
/* nrep device copies of `bytes` bytes by a plain streaming kernel; mean ms
 * per copy.  Reference point for calibrating HBM byte counters. */
int gfx_stream_copy(size_t bytes, int nrep, double *ms);

/* Sweep-kernel autotuner: on the first sweep of a lattice in a precision and
 * gauge, every variant is timed on a scratch copy of g and the fastest is
 * kept.  Variant v (0..31): workgroup size 64 if bit 0 else 256, star loop
 * unrolled if bit 1 else rolled (in the FP32 forms: software-pipelined, the
 * next direction's operands loaded while a direction's terms are formed),
 * non-temporal link loads if bit 2, and bits 3-4 the workgroup order: 0
 * lexicographic, 1 and 2 chunks of 4 and 8 superrows (x2 rows holding whole
 * workgroups) with x3 outermost, 3 chunks in x1 as well (variants whose
 * order does not fit the lattice are not timed; the PACKED64 and FIXED64
 * forms have rolled variants only).  on = 0 keeps variant 0 or the one set
 * by gfx_set_sweep_variant. */
int gfx_set_autotune(int on);
/* the chosen variant and the measured ms per half sweep of each of the 32
 * variants (0 when not timed) for precision GFX_PREC_FP64, GFX_PREC_MIXED
 * (the FP32 sweep), GFX_SWEEP_PACKED64 or GFX_SWEEP_PACKED32; GFX_ERR_STATE if not tuned yet */
int gfx_get_sweep_variant(int precision, int ndir, int *variant, int *block, int *rolled,
                          double *ms_per_variant);
int gfx_set_sweep_variant(int precision, int ndir, int variant);

/* ---- stochastic parts (MersenneTwister.cu, simulated_annealing.cu,
 *      heatbath_thermalization.cu extensions) ---- */
/* path of the twister parameter file (default "MersenneTwister.dat" in the
 * working directory, as MersenneTwister.cu:258) */
int gfx_set_mt_file(const char *path);
/* parameter records given directly: 4096 x {matrix_a, mask_b, mask_c, seed} */
int gfx_set_mt_params(const uint32_t *params);
/* copies of the gauss / uniform fields, the 4096 x 19 states and the params */
int gfx_get_mt_fields(float *gauss, float *unif, uint32_t *state, uint32_t *params);
/* numbers per twister of the gauss and uniform fields */
int gfx_mt_sizes(int *n_per_gauss, int *n_per_unif);
/* SA options: sampler 0 = the reference's rule, 1 = Kennedy-Pendleton;
 * nmicro microcanonical sweeps after each SA sweep; ndir 4 Landau / 3 Coulomb */
int gfx_set_sa(int sampler, int nmicro, int ndir);
/* n microcanonical sweeps of g (even + odd): in every SU(2) subgroup g is
 * reflected so that Re tr(g K) is unchanged (heatbath_thermalization.cu:318-338
 * rule applied to g) */
int gfx_micro_sweeps(int n, int ndir);
/* SA temperature of step i (simulated_annealing.cu:265-299) */
double gfx_sa_temperature(int i, double Tmin, double Tmax, double expo, int N);
/* n draws of the SU(2) subgroup sampler of the SA sweep (sampler 0: the
 * reference rule of simulated_annealing.cu:5-73, 1: Kennedy-Pendleton) for
 * w = (wnorm, 0, 0, 0) at inverse temperature beta, from the current twister
 * fields, which are refilled afterwards; a0[i] = alpha_0 (NaN where the
 * Kennedy-Pendleton trials all failed).  For distribution tests. */
int gfx_su2_sample(int sampler, double beta, double wnorm, int n, double *a0);
/* the same, all four components of each draw: alpha[4i..4i+3] = (alpha_0,
 * alpha_1, alpha_2, alpha_3) in the reference's component order
 * (simulated_annealing.cu:47-55), NaN where the Kennedy-Pendleton trials
 * all failed */
int gfx_su2_sample4(int sampler, double beta, double wnorm, int n, double *alpha);
/* nsweep link-heatbath sweeps of the Wilson action at beta (Kennedy-Pendleton
 * SU(2) subgroups, random numbers from the twisters) */
int gfx_heatbath_sweeps(int nsweep, double beta);

/* ================= batching form (no counterpart in the reference) =================
 * Many lattices gauge fixed together: Gribov copies of one configuration
 * (sharing its links) and configurations of different sizes.  Every half sweep
 * is one launch per precision over the lattices sweeping in it; a lattice
 * stops, or (mixed) moves from FP32 to FP64, at its own check; after a short
 * eager start (which also times the sweep variants and keeps the fastest),
 * `periods_per_graph` check periods of `checkinterval` iterations each are
 * captured in a HIP graph and replayed until some lattice changes state.
 * Each lattice follows overrelax_gauge's stop rule and records its own stop
 * iteration (the iterations it ran). */
int gfx_batch_create(void);
void gfx_batch_destroy(void);
int gfx_batch_size(void);
/* a lattice with its own links (4V su3) and initial g (V su3); returns its
 * index or a negative status */
int gfx_batch_add(int lx, int ly, int lz, int lt, gfx_su3 *gf, gfx_su3 *trafo);
/* a copy sharing the links of lattice `base`, initial g from `trafo` */
int gfx_batch_add_copy(int base, gfx_su3 *trafo);
/* a copy of `base` with a Haar-random initial g drawn from the twisters
 * (needs MersenneTwister.dat or gfx_set_mt_params) */
int gfx_batch_add_random_copy(int base);
/* precision GFX_PREC_FP64 or GFX_PREC_MIXED (each lattice: FP32 until its max
 * dAdA < switch_theta or its FP32 stall, then FP64); returns GFX_OK when every
 * lattice reached eps, else GFX_ERR_NOCONV */
int gfx_batch_run(int maxit, double eps, int checkinterval, int precision, double switch_theta,
                  int periods_per_graph);
int gfx_batch_result(int q, double *F, double *dAdA, double *maxdAdA, int *iter);
int gfx_batch_get_trafo(int q, gfx_su3 *trafo);
/* replaces lattice q's g (V su3, host layout): the start of the next
 * gfx_batch_run or gfx_batch_anneal */
int gfx_batch_set_trafo(int q, gfx_su3 *trafo);
int gfx_batch_stats(double *ms, int *graph_launches, int *iters_fp32, int *iters_fp64,
                    double *site_updates);
int gfx_batch_variant(int precision, int *variant, double *ms_per_variant);
/* Persistent launches for gfx_batch_run (on = 1; default 0, kept across
 * gfx_batch_create): while every lattice still sweeping is in one precision,
 * the check periods run inside one cooperative launch whose workgroups stay
 * resident, take the sweep blocks of all lattices from the workgroup map and
 * meet at a grid barrier between the half sweeps and around each check; the
 * launch returns when a lattice stops or leaves FP32, or after 32 periods.
 * Stretches with lattices in both precisions keep the graph replays.  Same
 * results as the graph replays.  GFX_ERR_DEVICE if the GPU has no
 * cooperative launch. */
int gfx_batch_set_persistent(int on);
/* persistent launches of the last gfx_batch_run */
int gfx_batch_persistent_launches(void);
/* simulated annealing of every lattice's g (simannealing_gauge per lattice,
 * simulated_annealing.cu:309-476): nsteps steps at gfx_sa_temperature(i,
 * Tmin, Tmax, expo, nsteps), each an even and an odd heatbath half sweep
 * (sampler 0: the reference rule, 1: Kennedy-Pendleton) with a twister refill
 * after each, then nmicro microcanonical sweeps; one launch per half sweep
 * over the batch, steps replayed as a HIP graph of steps_per_graph steps */
int gfx_batch_anneal(int nsteps, double Tmin, double Tmax, double expo, int nmicro, int sampler,
                     int steps_per_graph);
/* precision of gfx_batch_anneal: GFX_PREC_FP64 (default; the reference's SA
 * arithmetic, simulated_annealing.cu:5-257) or GFX_PREC_MIXED: links and g in
 * FP32 with FP32 arithmetic (the SU(2) sampling stays FP64) for the whole
 * schedule, g reunitarised into FP64 at the end.  Kept until changed (also
 * across gfx_batch_create). */
int gfx_batch_set_anneal_precision(int mode);
/* after gfx_batch_run: the lattice with the largest F among those sharing
 * the links of `base` (the Gribov copy to keep) */
int gfx_batch_best_copy(int base, int *q_best, double *F_best);

/* ================= no allocation (no counterpart in the reference) =================
 * For callers that cannot allocate device memory: code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc.  F, dAdA
 * and max dAdA (calc_functional's three numbers, overrelax.cu:610-671) of
 * caller-owned device fields in the reference's device layout (2-row
 * storage, su3manip.c:212-230, 281-299: U_mu(s) at gf[12 (4 s + mu) ..],
 * g(s) at trafo[12 s ..], 12 doubles each, s lexicographic), written to the
 * device array result[3] = {F, dAdA, max dAdA} by kernels launched on
 * `stream` (a hipStream_t; NULL the null stream) without waiting for them.
 * Nothing is allocated: the reduction's partials live in a static device
 * array of the library (1.5 MiB per GPU, held until the process ends), so at
 * most 16,777,216 sites (64^4) are accepted (GFX_ERR_SIZE beyond).  Calls
 * that overlap on one GPU share that array and race.  Independent of
 * init_gaugefixing; returns GFX_OK, GFX_ERR_GEOMETRY, GFX_ERR_SIZE or
 * GFX_ERR_DEVICE. */
int gfx_calc_functional_dev(const double *gf, const double *trafo, int lx, int ly, int lz, int lt,
                            double *result, void *stream);

/* last error as text */
const char *gfx_error_string(int code);

#ifdef __cplusplus
}
#endif
#endif
