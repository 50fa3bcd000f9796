// gfx_state.h -- device state of the single library instance (the reference
// keeps the same state in globals, global.h:128-170, cudagaugefix.cu:46-54).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "gfx_lattice.h"

namespace gfx {

// sweep forms: FP64 links and g, FP32 links and g, the forms on packed
// links (FP64 g on 24-bit high + 16-bit low parts: Coulomb and batch polish
// below the fixed-point polish's reach; FP32 g on the 21-bit links: the FP32
// phase of mixed runs), and the FP64 polish on 32-bit fixed-point links and g
enum { FORM_F64 = 0, FORM_F32 = 1, FORM_P64 = 2, FORM_P32 = 3, FORM_X64 = 4, NFORM = 5 };

struct State {
  int init = 0;  // 1 gauge fixing, 2 thermalization
  Geom geo{};
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;

  // FP64 fields: links 8 fields (parity*4+mu), trafo 3 fields: even, odd,
  // and the spare even field that a check sweep writes into (between calls
  // the even field is always field 0)
  double2* U64 = nullptr;
  double2* g64 = nullptr;
  // FP32 12-real copies for the mixed path (same field layout)
  float4* U32 = nullptr;
  float4* g32 = nullptr;
  bool U32_valid = false;
  // packed links (PackedLinks) of the mixed Landau run, made from U64;
  // valid while U32 is (both are invalidated together)
  PackedArgs pk{};
  bool packed_valid = false;
  // 32-bit fixed-point links (8 fields) and g (4 fields, slots as g64) of
  // the mixed Landau run's FP64 polish; Ufx valid while U32 is
  int4* Ufx = nullptr;
  int4* gfx = nullptr;
  bool fx_valid = false;
  // the FP64 sweeps of the Landau run in progress read Ufx / gfx
  bool fx_on = false;

  // fused-check partials: odd sweep [0, ncheck), even sweep [ncheck, 2 ncheck)
  double* cpart = nullptr;
  int ncheck = 0;

  // reductions
  double* partials = nullptr;
  int npartials = 0;
  double* results = nullptr;    // device, 4 per segment
  double* h_results = nullptr;  // pinned host mirror
  unsigned char* active = nullptr;

  double* staging = nullptr;  // AoS transfer chunk

  // parameters
  double omega = 1.68;  // overrelax.cu:416
  int precision = 0;
  double switch_theta = 1e-6;
  // the FP32 phase also ends when its best max dAdA has stayed below
  // stall_gate without halving for stall_iters iterations (rounding floor)
  int stall_iters = 300;
  double stall_gate = 1e-8;
  // a mixed Landau run sweeps on packed links (1: FP32 phase on their high
  // parts, FP64 polish on high + low parts) or on the FP32 and FP64 links
  // (0); packed_on: the Landau sweeps in progress read the packed links
  int packed_links = 1;
  bool packed_on = false;
  int verbose = 1;
  long mt_seed = -1;

  // Mersenne Twister fields (MersenneTwister.cu:55-68)
  uint32_t* mt_state = nullptr;   // 4096 x 19
  uint32_t* mt_params = nullptr;  // 4096 x 4 (matrix_a, mask_b, mask_c, seed)
  float* rnd_gauss = nullptr;
  float* rnd_unif = nullptr;
  int n_per_gauss = 0, n_per_unif = 0;
  // a second set of random fields and a stream of its own for the twister
  // refills, so that the refill for the next link-heatbath launch runs
  // beside the current one (same numbers, same order)
  float* rnd_gauss2 = nullptr;
  float* rnd_unif2 = nullptr;
  int n2_gauss = 0, n2_unif = 0;
  hipStream_t mt_stream = nullptr;
  hipEvent_t ev_used[2] = {nullptr, nullptr};    // a launch finished reading buffer k
  hipEvent_t ev_filled[2] = {nullptr, nullptr};  // buffer k refilled
  // workgroup order (tiling of bmap[0]) of the link heatbath for this
  // lattice, picked by timing on its first sweep (-1: not yet)
  int hb_tiling = -1;
  // the same for the simulated-annealing and microcanonical sweeps of g,
  // picked by timing the first SA launches of the first annealing run
  int sa_tiling = -1;

  // sweep kernel variant per [form: FORM_*][ndir
  // 3/4], picked by the autotuner on first use for the current lattice
  int variant[NFORM][2] = {};
  bool tuned[NFORM][2] = {};
  float tune_ms[NFORM][2][32] = {};
  int autotune = 1;
  // workgroup maps per [block 256/64][tiling] (device), and whether the
  // tiling applies to this lattice
  int* bmap[2][4] = {};
  bool tiling_ok[2][4] = {};
  // Coulomb: per precision, a map of the blocks of the slices still being
  // fixed (nblocks entries; 0 = use the full map)
  int* smap[2] = {nullptr, nullptr};
  int smap_n[2] = {0, 0};
  std::vector<int> smap_h[2];

  // last run statistics
  double run_ms = 0.0;
  int run_iters32 = 0, run_iters64 = 0;

  // launch times of the half sweeps without a check inside real runs
  // (gfx_set_sweep_sampling): each run of consecutive such launches of one
  // form (between two checks) is a segment bracketed by an event pair on the
  // library stream, at most SAMPLE_MAX segments per run
  static constexpr int SAMPLE_MAX = 2048;
  int sample_on = 0;
  int sample_n = 0;  // closed segments
  bool sample_open = false;
  hipEvent_t* sample_ev = nullptr;  // 2 * SAMPLE_MAX: start, end
  int sample_form[SAMPLE_MAX] = {};
  int sample_cnt[SAMPLE_MAX] = {};  // launches in the segment
};

extern State g_st;

int hip_check(hipError_t e, const char* what);
// one functional evaluation into s.h_results for nseg segments, the even
// field in slot `even` and the odd field in slot `odd`
int functional_now(State& s, bool fp32, int ndir, int nseg, double eps, int even = 0, int odd = 1);
Geom make_geom(int l0, int l1, int l2, int l3);
bool geometry_ok(int l0, int l1, int l2, int l3);
int ensure_common(int need_links32);
int dev_alloc(void** p, size_t bytes, const char* what);
size_t field_bytes64(const Geom& g);
size_t field_bytes32(const Geom& g);
size_t field_bytes_fx(const Geom& g);
int upload_aos_geo(State& s, const Geom& geo, const double* host, double2* dst, int nmat);
// twisters (gfx_stoch_api.hip): seed if needed and size the fields; refill
int mt_prepare(State& s, int need_g, int need_u);
void mt_refill(State& s, int initialized);
// the next refill's numbers into the given fields on stream st; the second
// set of fields (rnd_gauss2, rnd_unif2), the twister stream and its events
void mt_refill_into(State& s, float* gauss, float* unif, hipStream_t st);
int mt_second_set(State& s);
int download_aos_geo(State& s, const Geom& geo, const double2* src, double* host, int nmat);

}  // namespace gfx
