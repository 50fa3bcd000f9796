// gfx_stoch_api.hip -- C ABI of the stochastic parts: Mersenne Twister
// fields (MersenneTwister.cu:236-330), simulated annealing
// (simulated_annealing.cu:265-476) and link-heatbath thermalization
// (heatbath_thermalization.cu:557-841).
#include <hip/hip_runtime.h>

#include <algorithm>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <vector>

#include "../../include/cudagfx.h"
#include "gfx_internal.h"
#include "gfx_state.h"

namespace gfx {
void launch_mt_generate(uint32_t* params, uint32_t* state, float* out, int nper, int initialized,
                        hipStream_t st);
void launch_box_muller(float* d, int nper, hipStream_t st);
void launch_sa_sweep(const double2* U, double2* g, const Geom& geo, int parity, int ndir,
                     double beta, int sampler, const float* gauss, const float* unif,
                     const int* bmap, hipStream_t st);
void launch_micro_sweep(const double2* U, double2* g, const Geom& geo, int parity, int ndir,
                        const int* bmap, hipStream_t st);
void launch_link_heatbath(double2* U, const Geom& geo, int parity, int nu, double beta,
                          const float* gauss, const float* unif, const int* bmap, hipStream_t st);
int kp_gauss_per_site();
int kp_unif_per_site();
void launch_haar_trafo(double2* g, const Geom& geo, const float* gauss, hipStream_t st);
void launch_su2_sample(int sampler, double beta, double wnorm, int n, const float* gauss,
                       const float* unif, double* out, hipStream_t st, bool all);

static const int MT_COUNT = 4096;

struct Stoch {
  std::vector<uint32_t> host_params;  // 4096 x 4 as loaded (seed fields from the file)
  bool have_params = false;
  char dat_path[1024] = "MersenneTwister.dat";
  int sa_sampler = 0;   // 0 reference rule, 1 Kennedy-Pendleton
  int sa_nmicro = 0;    // microcanonical sweeps after each SA sweep
  int sa_ndir = 4;
};
static Stoch g_rng;

// iDivUp / iAlignUp (MersenneTwister.cu:31-43)
static int n_per_rng(int n) {
  int q = (n % MT_COUNT) ? n / MT_COUNT + 1 : n / MT_COUNT;
  return (q % 2) ? q + 1 : q;
}

static int load_params() {
  if (g_rng.have_params) return GFX_OK;
  FILE* fd = fopen(g_rng.dat_path, "rb");
  if (!fd) {
    printf("initMTGPU(): failed to open %s\n", g_rng.dat_path);
    printf("FAILED\n");
    return GFX_ERR_MTDAT;
  }
  g_rng.host_params.assign(4 * MT_COUNT, 0u);
  const size_t got = fread(g_rng.host_params.data(), sizeof(uint32_t) * 4 * MT_COUNT, 1, fd);
  fclose(fd);
  if (got != 1) {
    printf("initMTGPU(): failed to load %s\n", g_rng.dat_path);
    printf("FAILED\n");
    return GFX_ERR_MTDAT;
  }
  g_rng.have_params = true;
  return GFX_OK;
}

// (re)size the random fields; keeps the twister state
static int mt_fields(State& s, int n_gauss, int n_unif) {
  const int ng = n_per_rng(n_gauss), nu = n_per_rng(n_unif);
  if (ng > s.n_per_gauss || !s.rnd_gauss) {
    hipFree(s.rnd_gauss);
    if (hipMalloc((void**)&s.rnd_gauss, sizeof(float) * (size_t)MT_COUNT * (ng > 0 ? ng : 2)) != hipSuccess) return GFX_ERR_ALLOC;
  }
  if (nu > s.n_per_unif || !s.rnd_unif) {
    hipFree(s.rnd_unif);
    if (hipMalloc((void**)&s.rnd_unif, sizeof(float) * (size_t)MT_COUNT * (nu > 0 ? nu : 2)) != hipSuccess) return GFX_ERR_ALLOC;
  }
  s.n_per_gauss = ng;
  s.n_per_unif = nu;
  return GFX_OK;
}

// RandomGPU(gauss) -> BoxMullerGPU -> RandomGPU(unif), sharing one state
// array (MersenneTwister.cu:272-277, 300-306)
void mt_refill(State& s, int initialized) {
  launch_mt_generate(s.mt_params, s.mt_state, s.rnd_gauss, s.n_per_gauss, initialized, s.stream);
  launch_box_muller(s.rnd_gauss, s.n_per_gauss, s.stream);
  launch_mt_generate(s.mt_params, s.mt_state, s.rnd_unif, s.n_per_unif, initialized, s.stream);
}

int mt_init_device(State& s, int n_gauss, int n_unif) {
  int rc = load_params();
  if (rc) return rc;
  if (!s.stream) {
    if (hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking) != hipSuccess) return GFX_ERR_DEVICE;
    hipEventCreate(&s.ev0);
    hipEventCreate(&s.ev1);
  }
  // seedMTGPU (MersenneTwister.cu:98-114): srand, then one rand() per twister
  srand(s.mt_seed >= 0 ? (unsigned)s.mt_seed : (unsigned)time(NULL));
  std::vector<uint32_t> p(g_rng.host_params);
  for (int i = 0; i < MT_COUNT; i++) p[4 * i + 3] = (uint32_t)rand();
  if (!s.mt_params && hipMalloc((void**)&s.mt_params, sizeof(uint32_t) * 4 * MT_COUNT) != hipSuccess) return GFX_ERR_ALLOC;
  if (!s.mt_state && hipMalloc((void**)&s.mt_state, sizeof(uint32_t) * 19 * MT_COUNT) != hipSuccess) return GFX_ERR_ALLOC;
  hipMemcpyAsync(s.mt_params, p.data(), sizeof(uint32_t) * 4 * MT_COUNT, hipMemcpyHostToDevice, s.stream);
  if ((rc = mt_fields(s, n_gauss, n_unif))) return rc;
  printf("No. of gauss random numbers: %d\n", MT_COUNT * s.n_per_gauss);
  printf("No. of unif. dist. random numbers: %d\n", MT_COUNT * s.n_per_unif);
  mt_refill(s, 0);
  return hip_check(hipStreamSynchronize(s.stream), "init_MT");
}

// twisters seeded and fields holding at least need_g normals and need_u
// uniforms (grown and refilled if they were smaller)
int mt_prepare(State& s, int need_g, int need_u) {
  if (!s.mt_state) return mt_init_device(s, need_g, need_u);
  if (MT_COUNT * s.n_per_gauss < need_g || MT_COUNT * s.n_per_unif < need_u) {
    const int rc = mt_fields(s, std::max(need_g, MT_COUNT * s.n_per_gauss),
                             std::max(need_u, MT_COUNT * s.n_per_unif));
    if (rc) return rc;
    mt_refill(s, 1);
  }
  return GFX_OK;
}

// the twister output of one refill into the given fields on stream st
// (the same kernels and state as mt_refill)
void mt_refill_into(State& s, float* gauss, float* unif, hipStream_t st) {
  launch_mt_generate(s.mt_params, s.mt_state, gauss, s.n_per_gauss, 1, st);
  launch_box_muller(gauss, s.n_per_gauss, st);
  launch_mt_generate(s.mt_params, s.mt_state, unif, s.n_per_unif, 1, st);
}

// the second set of fields (as large as the first), the refill stream and
// its events
int mt_second_set(State& s) {
  if (!s.mt_stream) {
    if (hipStreamCreateWithFlags(&s.mt_stream, hipStreamNonBlocking) != hipSuccess) return GFX_ERR_DEVICE;
    for (int k = 0; k < 2; k++) {
      if (hipEventCreateWithFlags(&s.ev_used[k], hipEventDisableTiming) != hipSuccess ||
          hipEventCreateWithFlags(&s.ev_filled[k], hipEventDisableTiming) != hipSuccess)
        return GFX_ERR_DEVICE;
    }
  }
  if (s.n2_gauss < s.n_per_gauss || !s.rnd_gauss2) {
    hipFree(s.rnd_gauss2);
    s.rnd_gauss2 = nullptr;
    if (hipMalloc((void**)&s.rnd_gauss2, sizeof(float) * (size_t)MT_COUNT * s.n_per_gauss) != hipSuccess) return GFX_ERR_ALLOC;
    s.n2_gauss = s.n_per_gauss;
  }
  if (s.n2_unif < s.n_per_unif || !s.rnd_unif2) {
    hipFree(s.rnd_unif2);
    s.rnd_unif2 = nullptr;
    if (hipMalloc((void**)&s.rnd_unif2, sizeof(float) * (size_t)MT_COUNT * s.n_per_unif) != hipSuccess) return GFX_ERR_ALLOC;
    s.n2_unif = s.n_per_unif;
  }
  return GFX_OK;
}

void mt_free(State& s) {
  if (s.mt_stream) hipStreamSynchronize(s.mt_stream);
  hipFree(s.rnd_gauss2);
  hipFree(s.rnd_unif2);
  s.rnd_gauss2 = s.rnd_unif2 = nullptr;
  s.n2_gauss = s.n2_unif = 0;
  hipFree(s.rnd_gauss);
  hipFree(s.rnd_unif);
  hipFree(s.mt_params);
  hipFree(s.mt_state);
  s.rnd_gauss = s.rnd_unif = nullptr;
  s.mt_params = s.mt_state = nullptr;
  s.n_per_gauss = s.n_per_unif = 0;
}

}  // namespace gfx

using namespace gfx;

extern "C" {

// parameter globals of the input grammar (read_input.l:29-46), defined by
// the host input parser of this library
typedef struct {
  double Tmin, Tmax, expo;
  int N, checkint;
} gfx_saparam_t;
typedef struct {
  int Nsweep, Nsave;
  double beta;
  int checkint, startcond;
} gfx_thermparam_t;
extern gfx_saparam_t saparam;
extern gfx_thermparam_t thermparam;

int gfx_set_mt_seed(long seed) {
  g_st.mt_seed = seed;
  return GFX_OK;
}

int gfx_set_mt_file(const char* path) {
  if (!path || strlen(path) >= sizeof(g_rng.dat_path)) return GFX_ERR_STATE;
  strcpy(g_rng.dat_path, path);
  g_rng.have_params = false;
  return GFX_OK;
}

// twister parameter records supplied directly: 4096 x {matrix_a, mask_b, mask_c, seed}
int gfx_set_mt_params(const uint32_t* params) {
  g_rng.host_params.assign(params, params + 4 * MT_COUNT);
  g_rng.have_params = true;
  return GFX_OK;
}

int init_MT(int n_gaussnumbers, int n_unifnumbers) {
  printf("Initializing MT random number generator...\n");
  return mt_init_device(g_st, n_gaussnumbers, n_unifnumbers);
}

void update_MT(void) {
  State& s = g_st;
  if (!s.mt_state) return;
  mt_refill(s, 1);
  hip_check(hipStreamSynchronize(s.stream), "update_MT");
}

void finalize_MT(void) { mt_free(g_st); }

// copies of the random fields and twister state, for inspection
int gfx_get_mt_fields(float* gauss, float* unif, uint32_t* state, uint32_t* params) {
  State& s = g_st;
  if (!s.mt_state) return GFX_ERR_STATE;
  // the library stream does not synchronise with the null stream
  if (hip_check(hipStreamSynchronize(s.stream), "MT fields")) return GFX_ERR_DEVICE;
  if ((gauss && hip_check(hipMemcpy(gauss, s.rnd_gauss, sizeof(float) * MT_COUNT * s.n_per_gauss, hipMemcpyDeviceToHost), "D2H")) ||
      (unif && hip_check(hipMemcpy(unif, s.rnd_unif, sizeof(float) * MT_COUNT * s.n_per_unif, hipMemcpyDeviceToHost), "D2H")) ||
      (state && hip_check(hipMemcpy(state, s.mt_state, sizeof(uint32_t) * 19 * MT_COUNT, hipMemcpyDeviceToHost), "D2H")) ||
      (params && hip_check(hipMemcpy(params, s.mt_params, sizeof(uint32_t) * 4 * MT_COUNT, hipMemcpyDeviceToHost), "D2H")))
    return GFX_ERR_DEVICE;
  return GFX_OK;
}

// the MT field sizes per twister (N_PER_RNG_GAUSS, N_PER_RNG_UNIF)
int gfx_mt_sizes(int* n_per_gauss, int* n_per_unif) {
  if (n_per_gauss) *n_per_gauss = g_st.n_per_gauss;
  if (n_per_unif) *n_per_unif = g_st.n_per_unif;
  return GFX_OK;
}

// sampler 0: the reference's rule; 1: Kennedy-Pendleton.  nmicro
// microcanonical sweeps follow every SA sweep.  ndir 4 Landau, 3 Coulomb.
int gfx_set_sa(int sampler, int nmicro, int ndir) {
  g_rng.sa_sampler = sampler ? 1 : 0;
  g_rng.sa_nmicro = nmicro < 0 ? 0 : nmicro;
  g_rng.sa_ndir = (ndir == 3) ? 3 : 4;
  return GFX_OK;
}

// set_sa_temperature (simulated_annealing.cu:265-299): the temperature of step i
double gfx_sa_temperature(int i, double Tmin, double Tmax, double expo, int N) {
  double a, temperature;
  if (Tmax == Tmin) {
    temperature = Tmax;
  } else if (expo == 0) {
    a = (double)(i) / (double)(N - 1);
    temperature = pow(Tmin / Tmax, a) * Tmax;
  } else if (expo == -1) {
    a = (Tmin - Tmax) / (double)(N - 1);
    temperature = (a * i + Tmax);
  } else {
    a = pow(Tmin, -expo) - pow(Tmax, -expo);
    a = a / (double)(N - 1);
    temperature = pow((a * i + pow(Tmax, -expo)), (-1.0 / expo));
  }
  return temperature;
}

void simannealing_gauge(void) {
  State& s = g_st;
  if (s.init != 1) return;
  const int ndir = g_rng.sa_ndir;
  const int sampler = g_rng.sa_sampler;
  const int nh = s.geo.nh;
  // random numbers per half sweep: reference rule 1 gauss + 4 unif per site
  const int need_g = (sampler == 0) ? nh : kp_gauss_per_site() * nh;
  const int need_u = (sampler == 0) ? 4 * nh : kp_unif_per_site() * nh;
  if (!s.mt_state) {
    if (mt_init_device(s, need_g, need_u)) return;
  } else if (MT_COUNT * s.n_per_gauss < need_g || MT_COUNT * s.n_per_unif < need_u) {
    if (mt_fields(s, need_g, need_u)) return;
    mt_refill(s, 1);
  }
  if (mt_second_set(s)) return;
  const auto t0 = std::chrono::steady_clock::now();
  const int ck = saparam.checkint > 0 ? saparam.checkint : 1;
  // SA launch k reads fields[k % 2]; the refill that launch k+1 reads is made
  // into fields[(k+1) % 2] on the twister stream beside launch k and the
  // microcanonical sweeps after it (once launch k-1 has read those fields),
  // refills in order: the numbers of a refill after each launch
  float* fg[2] = {s.rnd_gauss, s.rnd_gauss2};
  float* fu[2] = {s.rnd_unif, s.rnd_unif2};
  hipEventRecord(s.ev_filled[0], s.stream);  // fields[0] hold the current numbers
  hipStreamWaitEvent(s.mt_stream, s.ev_filled[0], 0);
  // workgroup order: the first SA launches of the first run on this lattice
  // try each tiling that applies
  std::vector<int> tilings;
  if (s.sa_tiling < 0)
    for (int t = 0; t < TILINGS; t++)
      if (s.tiling_ok[0][t]) tilings.push_back(t);
  float best = 1e30f;
  int k = 0;
  for (int i = 0; i < saparam.N; i++) {
    const double T = gfx_sa_temperature(i, saparam.Tmin, saparam.Tmax, saparam.expo, saparam.N);
    const double beta = 1.0 / T;
    for (int p = 0; p < 2; p++, k++) {
      const int b = k & 1;
      const bool timing = k < (int)tilings.size();
      const int t = timing ? tilings[k] : (s.sa_tiling >= 0 ? s.sa_tiling : 0);
      if (k > 0) hipStreamWaitEvent(s.stream, s.ev_filled[b], 0);
      if (timing) hipEventRecord(s.ev0, s.stream);
      launch_sa_sweep(s.U64, s.g64, s.geo, p, ndir, beta, sampler, fg[b], fu[b], s.bmap[0][t], s.stream);
      if (timing) {
        hipEventRecord(s.ev1, s.stream);
        hipEventSynchronize(s.ev1);
        float ms = 0.f;
        hipEventElapsedTime(&ms, s.ev0, s.ev1);
        if (ms < best) {
          best = ms;
          s.sa_tiling = t;
        }
      }
      hipEventRecord(s.ev_used[b], s.stream);
      if (k > 0) hipStreamWaitEvent(s.mt_stream, s.ev_used[1 - b], 0);
      mt_refill_into(s, fg[1 - b], fu[1 - b], s.mt_stream);
      hipEventRecord(s.ev_filled[1 - b], s.mt_stream);
    }
    const int* map = s.bmap[0][s.sa_tiling >= 0 ? s.sa_tiling : 0];
    for (int m = 0; m < g_rng.sa_nmicro; m++)
      for (int p = 0; p < 2; p++) launch_micro_sweep(s.U64, s.g64, s.geo, p, ndir, map, s.stream);
    if ((i % ck) == 0) {
      if (functional_now(s, false, ndir, 1, 0.0)) return;
      FUNC = s.h_results[0];
      DADA = s.h_results[1];
      maxDADA = s.h_results[2];
      if (s.verbose)
        printf("iter %6d:\t FUNC = %.16e \t dAdA = %.16e\t max(dAdA) = %.16e\n", i, FUNC, DADA, maxDADA);
    }
  }
  // the fields of the last refill become the current ones
  hipStreamWaitEvent(s.stream, s.ev_filled[k & 1], 0);
  if (k & 1) {
    std::swap(s.rnd_gauss, s.rnd_gauss2);
    std::swap(s.rnd_unif, s.rnd_unif2);
  }
  hipStreamSynchronize(s.mt_stream);
  hipStreamSynchronize(s.stream);
  const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (s.verbose) printf("SA finished after %f sec\n", sec);
}

int gfx_batch_add_copy(int base, gfx_su3* trafo);
int gfx_batch_trafo_field(int q, void** g64, int dims[4]);

// A Gribov copy of lattice `base` whose initial g is a Haar-random gauge
// transformation drawn from the twister streams (12 normals per site).
// Fixing U from g = h is fixing the transformed field U^h from g = 1.
int gfx_batch_add_random_copy(int base) {
  State& s = g_st;
  const int q = gfx_batch_add_copy(base, nullptr);
  if (q < 0) return q;
  void* gp = nullptr;
  int d[4];
  gfx_batch_trafo_field(q, &gp, d);
  const Geom geo = make_geom(d[0], d[1], d[2], d[3]);
  const int need = 12 * 2 * geo.nh;
  if (!s.mt_state) {
    const int rc = mt_init_device(s, need, 2 * MT_COUNT);
    if (rc) return -rc;
  } else {
    if (MT_COUNT * s.n_per_gauss < need && mt_fields(s, need, MT_COUNT * s.n_per_unif)) return -GFX_ERR_ALLOC;
    mt_refill(s, 1);
  }
  launch_haar_trafo((double2*)gp, geo, s.rnd_gauss, s.stream);
  return hip_check(hipStreamSynchronize(s.stream), "haar") ? -GFX_ERR_DEVICE : q;
}

// n draws of the SU(2) sampler (0 reference rule, 1 Kennedy-Pendleton) for
// w = (wnorm, 0, 0, 0) at inverse temperature beta from the current twister
// fields (grown if needed), then one refill; a0[i] = alpha_0
static int su2_sample(int sampler, double beta, double wnorm, int n, double* out, bool all) {
  State& s = g_st;
  if (n <= 0 || !out) return GFX_ERR_STATE;
  const long ng = (long)(sampler ? kp_gauss_per_site() : 1) * n;
  const long nu = (long)(sampler ? kp_unif_per_site() : 4) * n;
  if (nu > (1L << 30)) return GFX_ERR_GEOMETRY;
  int rc = mt_prepare(s, (int)ng, (int)nu);
  if (rc) return rc;
  const size_t bytes = sizeof(double) * (size_t)n * (all ? 4 : 1);
  double* d = nullptr;
  if (hipMalloc((void**)&d, bytes) != hipSuccess) return GFX_ERR_ALLOC;
  launch_su2_sample(sampler ? 1 : 0, beta, wnorm, n, s.rnd_gauss, s.rnd_unif, d, s.stream, all);
  hipMemcpyAsync(out, d, bytes, hipMemcpyDeviceToHost, s.stream);
  mt_refill(s, 1);
  rc = hip_check(hipStreamSynchronize(s.stream), "su2 sample");
  hipFree(d);
  return rc;
}

int gfx_su2_sample(int sampler, double beta, double wnorm, int n, double* a0) {
  return su2_sample(sampler, beta, wnorm, n, a0, false);
}

// the same draws, all four quaternion components: alpha[4i..4i+3]
int gfx_su2_sample4(int sampler, double beta, double wnorm, int n, double* alpha) {
  return su2_sample(sampler, beta, wnorm, n, alpha, true);
}

// n microcanonical sweeps (even + odd) of g; F is unchanged by each
int gfx_micro_sweeps(int n, int ndir) {
  State& s = g_st;
  if (s.init != 1) return GFX_ERR_STATE;
  const int* map = s.bmap[0][s.sa_tiling >= 0 ? s.sa_tiling : 0];
  for (int k = 0; k < n; k++)
    for (int p = 0; p < 2; p++) launch_micro_sweep(s.U64, s.g64, s.geo, p, ndir == 3 ? 3 : 4, map, s.stream);
  return hip_check(hipStreamSynchronize(s.stream), "micro");
}

// one full heatbath sweep (4 directions x 2 parities) at beta_eff = beta/3
int gfx_heatbath_sweeps(int nsweep, double beta) {
  State& s = g_st;
  if (!s.init) return GFX_ERR_STATE;
  const int nh = s.geo.nh;
  const int need_g = kp_gauss_per_site() * nh, need_u = kp_unif_per_site() * nh;
  if (!s.mt_state) {
    int rc = mt_init_device(s, need_g, need_u);
    if (rc) return rc;
  } else if (MT_COUNT * s.n_per_gauss < need_g || MT_COUNT * s.n_per_unif < need_u) {
    if (mt_fields(s, need_g, need_u)) return GFX_ERR_ALLOC;
    mt_refill(s, 1);
  }
  int rc = mt_second_set(s);
  if (rc) return rc;
  // Launch k reads fields[k % 2]; the refill that launch k+1 reads is made
  // into fields[(k+1) % 2] on the twister stream while launch k runs (after
  // launch k-1 has finished reading those fields), refills in order, so
  // every launch gets the same numbers as with a refill after each launch.
  float* fg[2] = {s.rnd_gauss, s.rnd_gauss2};
  float* fu[2] = {s.rnd_unif, s.rnd_unif2};
  hipEventRecord(s.ev_filled[0], s.stream);  // fields[0] hold the current numbers
  hipStreamWaitEvent(s.mt_stream, s.ev_filled[0], 0);
  // workgroup order: the first launches of the first call on this lattice
  // try each tiling that applies (the result does not depend on it)
  std::vector<int> tilings;
  if (s.hb_tiling < 0)
    for (int t = 0; t < TILINGS; t++)
      if (s.tiling_ok[0][t]) tilings.push_back(t);
  float best = 1e30f;
  int bestt = 0;
  const int n = 8 * nsweep;
  for (int k = 0; k < n; k++) {
    const int nu = (k >> 1) & 3, p = k & 1, b = k & 1;
    const bool timing = k < (int)tilings.size();
    const int t = timing ? tilings[k] : (s.hb_tiling >= 0 ? s.hb_tiling : 0);
    if (k > 0) hipStreamWaitEvent(s.stream, s.ev_filled[b], 0);
    if (timing) hipEventRecord(s.ev0, s.stream);
    launch_link_heatbath(s.U64, s.geo, p, nu, beta / 3.0, fg[b], fu[b], s.bmap[0][t], s.stream);
    if (timing) {
      hipEventRecord(s.ev1, s.stream);
      hipEventSynchronize(s.ev1);
      float ms = 0.f;
      hipEventElapsedTime(&ms, s.ev0, s.ev1);
      if (ms < best) {
        best = ms;
        bestt = t;
      }
      if (k + 1 == (int)tilings.size()) s.hb_tiling = bestt;
    }
    hipEventRecord(s.ev_used[b], s.stream);
    // refill k+1 into the other fields once launch k-1 has read them
    if (k > 0) hipStreamWaitEvent(s.mt_stream, s.ev_used[1 - b], 0);
    mt_refill_into(s, fg[1 - b], fu[1 - b], s.mt_stream);
    hipEventRecord(s.ev_filled[1 - b], s.mt_stream);
  }
  // the fields of the last refill become the current ones
  hipStreamWaitEvent(s.stream, s.ev_filled[n & 1], 0);
  if (n & 1) {
    std::swap(s.rnd_gauss, s.rnd_gauss2);
    std::swap(s.rnd_unif, s.rnd_unif2);
  }
  s.U32_valid = false;
  if (hip_check(hipStreamSynchronize(s.mt_stream), "twister refill")) return GFX_ERR_DEVICE;
  return hip_check(hipStreamSynchronize(s.stream), "heatbath");
}

void thermalize_gauge(void) {
  State& s = g_st;
  if (!s.init) return;
  printf("Setting beta to %f\n", thermparam.beta);
  FILE* plaqfile = fopen("plaq.txt", "w");
  if (!plaqfile) fprintf(stderr, "Could not open file 'plaq.txt'\n");
  const auto t0 = std::chrono::steady_clock::now();
  const int ck = thermparam.checkint > 0 ? thermparam.checkint : 1;
  for (int i = 0; i < thermparam.Nsweep; i++) {
    if (gfx_heatbath_sweeps(1, thermparam.beta)) break;
    if ((i % ck) == 0) {
      const double plaq = calc_plaquette(0);
      if (plaqfile) fprintf(plaqfile, "%d %.16e\n", i, plaq);
      printf("therm %6d:\t PLAQ = %.16e\n", i, plaq);
    }
  }
  const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (plaqfile) fclose(plaqfile);
  printf("Thermalization finished after %f sec\n", sec);
}

}  // extern "C"
