// gfx_batch_api.hip -- the batching form: many lattices gauge fixed together
// (Gribov copies of one configuration, or configurations of different sizes)
// with one sweep launch per half sweep covering every lattice, device-side
// convergence flags, checks fused into the sweeps (see k_or_sweep), and the
// sweeps of check periods captured in a HIP graph that is replayed until
// every lattice has converged.
//
// No counterpart in the reference (it fixes one configuration per run); each
// lattice follows overrelax_gauge (overrelax.cu:677-823) exactly: checks at
// i % ck == 0 after sweep i, stop at max dAdA < eps, the stop iteration is
// recorded per lattice.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/cudagfx.h"
#include "gfx_internal.h"
#include "gfx_lattice.h"
#include "gfx_state.h"

namespace gfx {

// g64 / g32: three fields each (even, odd, spare even field of the checks);
// between phases the even field is field 0
struct BProbHost {
  Geom geo{};
  double2* U64 = nullptr;
  float4* U32 = nullptr;
  double2* g64 = nullptr;
  float4* g32 = nullptr;
  int ushare = 0;  // problem whose links this one uses (itself if it owns them)
  int stop_iter[2] = {-1, -1};  // per phase (FP32 bulk, FP64)
};

struct Batch {
  std::vector<BProbHost> probs;
  BatchProblem* d_probs = nullptr;
  int2* d_bmap[2][TILINGS] = {};  // per [block 256/64][tiling]
  int nbmap[2] = {0, 0};
  int* d_poff[2] = {nullptr, nullptr};  // check-partial offsets per problem
  double* d_podd = nullptr;   // check partials of the odd sweeps
  double* d_peven = nullptr;  // and of the even sweeps
  int* d_roff = nullptr;      // per problem: first site of its random numbers
  int nh_total = 0;           // sites per parity over all problems
  double* d_best = nullptr;   // per problem: best max dAdA of the phase
  int* d_stall = nullptr;     // checks since it last halved
  int* d_active = nullptr;
  int* h_active = nullptr;
  double* d_results = nullptr;
  double* h_results = nullptr;
  int* d_iter = nullptr;
  bool built = false;
  int variant[2] = {0, 0};
  bool tuned[2] = {false, false};
  float tune_ms[2][SWEEP_VARIANTS] = {};
  // statistics of the last run
  double run_ms = 0.0;
  int graph_launches = 0, iters32 = 0, iters64 = 0;
  double site_updates = 0.0;
};
static Batch g_b;

static const int BLOCKS[2] = {256, 64};

static int vblock_index(int variant) { return block_index(sweep_variant_block(variant)); }

static void batch_free_maps(Batch& b) {
  for (int k = 0; k < 2; k++) {
    for (int t = 0; t < TILINGS; t++) {
      hipFree(b.d_bmap[k][t]);
      b.d_bmap[k][t] = nullptr;
    }
    hipFree(b.d_poff[k]);
    b.d_poff[k] = nullptr;
  }
  hipFree(b.d_probs);
  hipFree(b.d_podd);
  hipFree(b.d_peven);
  hipFree(b.d_best);
  hipFree(b.d_roff);
  hipFree(b.d_stall);
  hipFree(b.d_active);
  hipFree(b.d_results);
  hipFree(b.d_iter);
  if (b.h_active) hipHostFree(b.h_active);
  if (b.h_results) hipHostFree(b.h_results);
  b.d_probs = nullptr;
  b.d_podd = nullptr;
  b.d_peven = nullptr;
  b.d_best = nullptr;
  b.d_roff = nullptr;
  b.d_stall = nullptr;
  b.d_active = nullptr;
  b.d_results = nullptr;
  b.d_iter = nullptr;
  b.h_active = nullptr;
  b.h_results = nullptr;
  b.built = false;
}

static void batch_free(Batch& b) {
  if (g_st.stream) hipStreamSynchronize(g_st.stream);
  for (auto& p : b.probs) {
    hipFree(p.g64);
    hipFree(p.g32);
    if (&p - &b.probs[0] == p.ushare) {
      hipFree(p.U64);
      hipFree(p.U32);
    }
  }
  b.probs.clear();
  batch_free_maps(b);
  for (int k = 0; k < 2; k++) b.tuned[k] = false;
}

// block maps: every lattice's blocks in the workgroup order of the tiling
// (block_order); Gribov copies sharing links are interleaved so that the
// blocks covering the same sites of every copy run together and share the
// link lines in L2 / Infinity Cache; lattices with their own links follow one
// another; the whole list is dealt to the XCDs in contiguous eighths.
static int batch_build(Batch& b) {
  batch_free_maps(b);
  const int n = (int)b.probs.size();
  for (int k = 0; k < 2; k++) {
    for (int t = 0; t < TILINGS; t++) {
      std::vector<int2> list;
      std::vector<bool> done(n, false);
      for (int q0 = 0; q0 < n; q0++) {
        if (done[q0]) continue;
        std::vector<int> group;
        for (int q = q0; q < n; q++)
          if (b.probs[q].ushare == b.probs[q0].ushare) {
            group.push_back(q);
            done[q] = true;
          }
        std::vector<int> order;
        block_order(b.probs[q0].geo, BLOCKS[k], t, order);
        for (int blk : order)
          for (int q : group) list.push_back(make_int2(q, blk));
      }
      std::vector<int2> map;
      deal_to_xcds(list, map);
      b.nbmap[k] = (int)map.size();
      if (dev_alloc((void**)&b.d_bmap[k][t], sizeof(int2) * map.size(), "batch map")) return GFX_ERR_ALLOC;
      hipMemcpy(b.d_bmap[k][t], map.data(), sizeof(int2) * map.size(), hipMemcpyHostToDevice);
    }
    std::vector<int> poff(n + 1, 0);
    for (int q = 0; q < n; q++) poff[q + 1] = poff[q] + (b.probs[q].geo.nh + BLOCKS[k] - 1) / BLOCKS[k];
    if (dev_alloc((void**)&b.d_poff[k], sizeof(int) * (n + 1), "batch offsets")) return GFX_ERR_ALLOC;
    hipMemcpy(b.d_poff[k], poff.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice);
  }
  std::vector<int> roff(n + 1, 0);
  for (int q = 0; q < n; q++) roff[q + 1] = roff[q] + b.probs[q].geo.nh;
  b.nh_total = roff[n];
  if (dev_alloc((void**)&b.d_roff, sizeof(int) * (n + 1), "batch random offsets")) return GFX_ERR_ALLOC;
  hipMemcpy(b.d_roff, roff.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice);
  // partial slots: as many as blocks of the smallest workgroup
  size_t npart = 0;
  for (int q = 0; q < n; q++) npart += (b.probs[q].geo.nh + 63) / 64;
  if (dev_alloc((void**)&b.d_podd, sizeof(double) * 3 * npart, "batch partials") ||
      dev_alloc((void**)&b.d_peven, sizeof(double) * 3 * npart, "batch partials") ||
      dev_alloc((void**)&b.d_probs, sizeof(BatchProblem) * n, "batch problems") ||
      dev_alloc((void**)&b.d_best, sizeof(double) * n, "batch best") ||
      dev_alloc((void**)&b.d_stall, sizeof(int) * n, "batch stall") ||
      dev_alloc((void**)&b.d_active, sizeof(int) * n, "batch active") ||
      dev_alloc((void**)&b.d_results, sizeof(double) * 4 * n, "batch results") ||
      dev_alloc((void**)&b.d_iter, sizeof(int), "batch iter"))
    return GFX_ERR_ALLOC;
  if (hipHostMalloc((void**)&b.h_active, sizeof(int) * n, 0) != hipSuccess ||
      hipHostMalloc((void**)&b.h_results, sizeof(double) * 4 * n, 0) != hipSuccess)
    return GFX_ERR_ALLOC;
  b.built = true;
  return GFX_OK;
}

template <typename V>
static V* bfield(V* g, const Geom& geo, int k) {
  return g + (size_t)k * Store<V>::NPLANE * geo.stride;
}

// canonical field pointers of the phase's precision: even = field 0, odd =
// field 1, spare = field 2
static void set_phase_pointers(Batch& b, bool fp32) {
  std::vector<BatchProblem> h(b.probs.size());
  for (size_t q = 0; q < b.probs.size(); q++) {
    const BProbHost& p = b.probs[q];
    const BProbHost& o = b.probs[p.ushare];
    h[q].geo = p.geo;
    h[q].U = fp32 ? (const void*)o.U32 : (const void*)o.U64;
    for (int k = 0; k < 2; k++)
      h[q].g[k] = fp32 ? (void*)bfield(p.g32, p.geo, k) : (void*)bfield(p.g64, p.geo, k);
    h[q].galt = fp32 ? (void*)bfield(p.g32, p.geo, 2) : (void*)bfield(p.g64, p.geo, 2);
  }
  hipStreamSynchronize(g_st.stream);
  hipMemcpy(b.d_probs, h.data(), sizeof(BatchProblem) * h.size(), hipMemcpyHostToDevice);
}

// after a phase: copy each even field the checks left in the spare slot back
// to field 0
static void batch_canonical(Batch& b, bool fp32) {
  State& s = g_st;
  std::vector<BatchProblem> h(b.probs.size());
  hipStreamSynchronize(s.stream);
  hipMemcpy(h.data(), b.d_probs, sizeof(BatchProblem) * h.size(), hipMemcpyDeviceToHost);
  for (size_t q = 0; q < b.probs.size(); q++) {
    const BProbHost& p = b.probs[q];
    void* f0 = fp32 ? (void*)p.g32 : (void*)p.g64;
    if (h[q].g[0] != f0)
      hipMemcpyAsync(f0, h[q].g[0], fp32 ? field_bytes32(p.geo) : field_bytes64(p.geo),
                     hipMemcpyDeviceToDevice, s.stream);
  }
  hipStreamSynchronize(s.stream);
}

static int ensure_fp32(Batch& b) {
  State& s = g_st;
  for (size_t q = 0; q < b.probs.size(); q++) {
    BProbHost& p = b.probs[q];
    if ((int)q == p.ushare) {
      if (!p.U32 && dev_alloc((void**)&p.U32, 8 * field_bytes32(p.geo), "batch FP32 links")) return GFX_ERR_ALLOC;
      launch_f64_to_f32(p.U64, p.U32, p.geo, 8, s.stream);
    }
    if (!p.g32 && dev_alloc((void**)&p.g32, 3 * field_bytes32(p.geo), "batch FP32 trafo")) return GFX_ERR_ALLOC;
    launch_f64_to_f32(p.g64, p.g32, p.geo, 2, s.stream);
  }
  return GFX_OK;
}

// one half sweep of every active lattice; chk as in launch_batch_sweep
static void batch_half(Batch& b, bool fp32, int parity, int variant, int chk) {
  State& s = g_st;
  const int k = vblock_index(variant);
  launch_batch_sweep(fp32, b.d_probs, b.d_bmap[k][sweep_variant_tiling(variant)], b.nbmap[k],
                     b.d_active, b.d_poff[k], parity, s.omega, variant, chk,
                     chk == 2 ? b.d_peven : b.d_podd, s.stream);
}

// the end of a check at iteration *d_iter: the even sweep of the next
// iteration carrying the check, then the per-lattice grid pass
static void batch_check_tail(Batch& b, bool fp32, int variant, double thr, int maxit,
                             int stall_checks) {
  State& s = g_st;
  batch_half(b, fp32, 0, variant, 2);
  launch_batch_finalize(b.d_podd, b.d_peven, b.d_poff[vblock_index(variant)], b.d_probs,
                        (int)b.probs.size(), 4, thr, b.d_active, b.d_results, b.d_iter, maxit,
                        b.d_best, b.d_stall, stall_checks, s.stall_gate, s.stream);
}

static int count_active(Batch& b) {
  State& s = g_st;
  hipMemcpyAsync(b.h_active, b.d_active, sizeof(int) * b.probs.size(), hipMemcpyDeviceToHost, s.stream);
  if (hip_check(hipStreamSynchronize(s.stream), "batch flags")) return -1;
  int n = 0;
  for (size_t q = 0; q < b.probs.size(); q++) n += b.h_active[q] ? 1 : 0;
  return n;
}

// Runs one phase (FP32 bulk or FP64) from global iteration *it until every
// lattice has stopped (max dAdA < thr, or an FP32 stall) or maxit is reached.
// The first iterations run eagerly (the sweep variants are timed on them when
// not tuned yet) up to the iteration after a check; then graphs of `periods`
// check periods are replayed while they end before maxit, and the rest runs
// eagerly again.
static int batch_phase(Batch& b, bool fp32, double thr, int ck, int maxit, int periods, int* it,
                       int stall_checks) {
  State& s = g_st;
  const int n = (int)b.probs.size();
  const int pi = fp32 ? 1 : 0;
  set_phase_pointers(b, fp32);
  launch_fill_i32(b.d_active, n, 1, s.stream);
  launch_fill_i32(b.d_stall, n, 0, s.stream);
  launch_fill_f64(b.d_best, n, 1e300, s.stream);
  int i = *it;
  bool even_done = false;  // the even half of iteration i ran with the last check
  int tune_v = b.tuned[pi] ? -1 : 0;
  int rc = GFX_OK;
  auto count_iter = [&]() {
    if (fp32) b.iters32++; else b.iters64++;
  };
  // one eager iteration; false when every lattice has stopped
  auto eager = [&](int v, bool timed) {
    const bool chk = (i % ck) == 0;
    const bool both = !even_done;
    if (timed && both) hipEventRecord(s.ev0, s.stream);
    if (both) batch_half(b, fp32, 0, v, 0);
    if (timed && !both) hipEventRecord(s.ev0, s.stream);
    batch_half(b, fp32, 1, v, chk ? 1 : 0);
    if (timed) {
      hipEventRecord(s.ev1, s.stream);
      hipEventSynchronize(s.ev1);
      float t = 0.f;
      hipEventElapsedTime(&t, s.ev0, s.ev1);
      b.tune_ms[pi][tune_v] = both ? 0.5f * t : t;
    }
    even_done = false;
    count_iter();
    bool go_on = true;
    if (chk) {
      launch_fill_i32(b.d_iter, 1, i, s.stream);
      batch_check_tail(b, fp32, v, thr, maxit, stall_checks);
      even_done = true;
      go_on = count_active(b) > 0;
    }
    i++;
    return go_on;
  };
  bool go_on = true;
  while (go_on && i < maxit) {
    int v = b.variant[pi];
    bool timed = false;
    if (tune_v >= 0) {
      // time full non-check iterations when there are any (ck >= 3), else
      // the odd half sweeps
      const bool chk = (i % ck) == 0;
      timed = (ck >= 3) ? (!even_done && !chk) : true;
      if (timed) v = tune_v;
    }
    go_on = eager(v, timed);
    if (timed && ++tune_v == SWEEP_VARIANTS) {
      int best = 0;
      for (int k = 1; k < SWEEP_VARIANTS; k++)
        if (b.tune_ms[pi][k] < b.tune_ms[pi][best]) best = k;
      b.variant[pi] = best;
      b.tuned[pi] = true;
      tune_v = -1;
    }
    if (tune_v < 0 && even_done) break;  // just after a check: periods start here
  }
  if (go_on && i + periods * ck <= maxit) {
    // graph of `periods` check periods over iterations i .. i + ck - 1, ...:
    // the even half of the first iteration ran with the previous check
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    const int v = b.variant[pi];
    if (hip_check(hipStreamBeginCapture(s.stream, hipStreamCaptureModeThreadLocal), "capture")) return GFX_ERR_DEVICE;
    for (int per = 0; per < periods; per++) {
      for (int j = 0; j < ck; j++) {
        if (j > 0) batch_half(b, fp32, 0, v, 0);
        batch_half(b, fp32, 1, v, j == ck - 1 ? 1 : 0);
      }
      batch_check_tail(b, fp32, v, thr, maxit, stall_checks);
      launch_batch_advance(b.d_iter, ck, s.stream);
    }
    if (hip_check(hipStreamEndCapture(s.stream, &graph), "end capture")) return GFX_ERR_DEVICE;
    if (hip_check(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "instantiate")) {
      hipGraphDestroy(graph);
      return GFX_ERR_DEVICE;
    }
    launch_fill_i32(b.d_iter, 1, i + ck - 1, s.stream);
    while (go_on && i + periods * ck <= maxit) {
      if (hip_check(hipGraphLaunch(exec, s.stream), "graph launch")) {
        rc = GFX_ERR_DEVICE;
        break;
      }
      b.graph_launches++;
      i += periods * ck;
      if (fp32) b.iters32 += periods * ck; else b.iters64 += periods * ck;
      const int na = count_active(b);
      if (na < 0) rc = GFX_ERR_DEVICE;
      go_on = na > 0;
    }
    hipGraphExecDestroy(exec);
    hipGraphDestroy(graph);
  }
  // tail: the iterations left before maxit
  while (rc == GFX_OK && go_on && i < maxit) go_on = eager(b.variant[pi], false);
  batch_canonical(b, fp32);
  *it = i;
  return rc;
}

}  // namespace gfx

using namespace gfx;

extern "C" {

int gfx_batch_create(void) {
  batch_free(g_b);
  return ensure_common(0);
}

void gfx_batch_destroy(void) { batch_free(g_b); }

int gfx_batch_size(void) { return (int)g_b.probs.size(); }

// a lattice with its own links; returns its index (>= 0) or -error
int gfx_batch_add(int l0, int l1, int l2, int l3, gfx_su3* gf, gfx_su3* trafo) {
  State& s = g_st;
  if (ensure_common(0)) return -GFX_ERR_DEVICE;
  if (!geometry_ok(l0, l1, l2, l3)) return -GFX_ERR_GEOMETRY;
  BProbHost p;
  p.geo = make_geom(l0, l1, l2, l3);
  p.ushare = (int)g_b.probs.size();
  if (dev_alloc((void**)&p.U64, 8 * field_bytes64(p.geo), "batch links")) return -GFX_ERR_ALLOC;
  if (dev_alloc((void**)&p.g64, 3 * field_bytes64(p.geo), "batch trafo")) {
    hipFree(p.U64);
    return -GFX_ERR_ALLOC;
  }
  if (upload_aos_geo(s, p.geo, (const double*)gf, p.U64, 4) ||
      upload_aos_geo(s, p.geo, (const double*)trafo, p.g64, 1)) {
    hipFree(p.U64);
    hipFree(p.g64);
    return -GFX_ERR_DEVICE;
  }
  g_b.probs.push_back(p);
  g_b.built = false;
  return p.ushare;
}

// a Gribov copy: the links of lattice `base`, its own initial g (V su3)
int gfx_batch_add_copy(int base, gfx_su3* trafo) {
  State& s = g_st;
  if (base < 0 || base >= (int)g_b.probs.size()) return -GFX_ERR_STATE;
  BProbHost p;
  const BProbHost& o = g_b.probs[base];
  p.geo = o.geo;
  p.ushare = o.ushare;
  p.U64 = g_b.probs[o.ushare].U64;
  if (dev_alloc((void**)&p.g64, 3 * field_bytes64(p.geo), "batch trafo")) return -GFX_ERR_ALLOC;
  if (trafo && upload_aos_geo(s, p.geo, (const double*)trafo, p.g64, 1)) {
    hipFree(p.g64);
    return -GFX_ERR_DEVICE;
  }
  g_b.probs.push_back(p);
  g_b.built = false;
  return (int)g_b.probs.size() - 1;
}

// device pointer and geometry of lattice q's FP64 g (used by the random copy initialiser)
int gfx_batch_trafo_field(int q, void** g64, int dims[4]) {
  if (q < 0 || q >= (int)g_b.probs.size()) return GFX_ERR_STATE;
  if (g64) *g64 = g_b.probs[q].g64;
  if (dims)
    for (int k = 0; k < 4; k++) dims[k] = g_b.probs[q].geo.L[k];
  return GFX_OK;
}

int gfx_batch_run(int maxit, double eps, int checkinterval, int precision, double switch_theta,
                  int periods_per_graph) {
  State& s = g_st;
  Batch& b = g_b;
  if (b.probs.empty()) return GFX_ERR_STATE;
  if (checkinterval < 1) checkinterval = 1;
  if (periods_per_graph < 1) periods_per_graph = 1;
  if (!b.built) {
    int rc = batch_build(b);
    if (rc) return rc;
  }
  const bool mixed = (precision == GFX_PREC_MIXED);
  if (mixed && ensure_fp32(b)) return GFX_ERR_ALLOC;
  b.graph_launches = b.iters32 = b.iters64 = 0;
  for (auto& p : b.probs) p.stop_iter[0] = p.stop_iter[1] = -1;
  hipStreamSynchronize(s.stream);
  const auto t0 = std::chrono::steady_clock::now();
  int it = 0, rc = GFX_OK;
  if (mixed) {
    launch_fill_f64(b.d_results, 4 * (int)b.probs.size(), -1.0, s.stream);
    rc = batch_phase(b, true, std::max(switch_theta > 0 ? switch_theta : s.switch_theta, eps),
                     checkinterval, maxit, periods_per_graph, &it,
                     (s.stall_iters + checkinterval - 1) / checkinterval);
    hipMemcpyAsync(b.h_results, b.d_results, sizeof(double) * 4 * b.probs.size(), hipMemcpyDeviceToHost, s.stream);
    hipStreamSynchronize(s.stream);
    for (size_t q = 0; q < b.probs.size(); q++) {
      b.probs[q].stop_iter[0] = (int)b.h_results[4 * q + 3];
      launch_g32_to_g64(b.probs[q].g32, b.probs[q].g64, b.probs[q].geo, s.stream);
    }
  }
  const int it_fp64_start = it;
  if (rc == GFX_OK) {
    launch_fill_f64(b.d_results, 4 * (int)b.probs.size(), -1.0, s.stream);
    rc = batch_phase(b, false, eps, checkinterval, maxit, periods_per_graph, &it, 0);
  }
  hipMemcpyAsync(b.h_results, b.d_results, sizeof(double) * 4 * b.probs.size(), hipMemcpyDeviceToHost, s.stream);
  hipStreamSynchronize(s.stream);
  b.run_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  int nconv = 0;
  b.site_updates = 0.0;
  for (size_t q = 0; q < b.probs.size(); q++) {
    BProbHost& p = b.probs[q];
    p.stop_iter[1] = (int)b.h_results[4 * q + 3];
    if (p.stop_iter[1] >= 0) nconv++;
    // iterations this lattice was updated in each phase
    const int n32 = mixed ? (p.stop_iter[0] >= 0 ? p.stop_iter[0] + 1 : it_fp64_start) : 0;
    const int n64 = (p.stop_iter[1] >= 0 ? p.stop_iter[1] + 1 : it) - it_fp64_start;
    b.site_updates += 2.0 * p.geo.nh * (double)(n32 + n64);
  }
  if (rc != GFX_OK) return rc;
  return nconv == (int)b.probs.size() ? GFX_OK : GFX_ERR_NOCONV;
}

// Simulated annealing of every lattice's g (simulated_annealing.cu:309-476
// per lattice): nsteps steps of an even and an odd heatbath half sweep at
// T(i) = gfx_sa_temperature(i, Tmin, Tmax, expo, nsteps), each half sweep
// followed by a refill of the twister fields, then nmicro microcanonical
// sweeps.  One launch per half sweep covers the batch; the random numbers of
// lattice q start at its site offset.  Steps run as replays of a graph of
// steps_per_graph steps, the temperature schedule and step counter on the
// device.
int gfx_batch_anneal(int nsteps, double Tmin, double Tmax, double expo, int nmicro, int sampler,
                     int steps_per_graph) {
  State& s = g_st;
  Batch& b = g_b;
  if (b.probs.empty() || nsteps < 0) return GFX_ERR_STATE;
  if (steps_per_graph < 1) steps_per_graph = 1;
  if (!b.built) {
    int rc = batch_build(b);
    if (rc) return rc;
  }
  sampler = sampler ? 1 : 0;
  nmicro = nmicro < 0 ? 0 : nmicro;
  const long ng = (long)(sampler ? kp_gauss_per_site() : 1) * b.nh_total;
  const long nu = (long)(sampler ? kp_unif_per_site() : 4) * b.nh_total;
  if (nu > (1L << 31) - 4096 * 2) return GFX_ERR_GEOMETRY;
  int rc = mt_prepare(s, (int)ng, (int)nu);
  if (rc) return rc;
  set_phase_pointers(b, false);
  std::vector<double> betas(nsteps > 0 ? nsteps : 1);
  for (int i = 0; i < nsteps; i++) betas[i] = 1.0 / gfx_sa_temperature(i, Tmin, Tmax, expo, nsteps);
  double* d_betas = nullptr;
  int* d_step = nullptr;
  if (dev_alloc((void**)&d_betas, sizeof(double) * betas.size(), "SA schedule") ||
      dev_alloc((void**)&d_step, sizeof(int), "SA step")) {
    hipFree(d_betas);
    return GFX_ERR_ALLOC;
  }
  hipMemcpy(d_betas, betas.data(), sizeof(double) * betas.size(), hipMemcpyHostToDevice);
  launch_fill_i32(d_step, 1, 0, s.stream);
  const int k = block_index(sa_block_size());
  const int2* map = b.d_bmap[k][0];
  const int nblk = b.nbmap[k];
  auto one_step = [&]() {
    for (int p = 0; p < 2; p++) {
      launch_batch_sa_sweep(b.d_probs, map, nblk, b.d_roff, p, d_betas, d_step, sampler,
                            s.rnd_gauss, s.rnd_unif, s.stream);
      mt_refill(s, 1);
    }
    for (int m = 0; m < nmicro; m++)
      for (int p = 0; p < 2; p++) launch_batch_micro_sweep(b.d_probs, map, nblk, p, s.stream);
    launch_batch_advance(d_step, 1, s.stream);
  };
  const auto t0 = std::chrono::steady_clock::now();
  int done = 0;
  if (nsteps >= steps_per_graph) {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    if (hip_check(hipStreamBeginCapture(s.stream, hipStreamCaptureModeThreadLocal), "capture")) rc = GFX_ERR_DEVICE;
    if (rc == GFX_OK) {
      for (int k2 = 0; k2 < steps_per_graph; k2++) one_step();
      if (hip_check(hipStreamEndCapture(s.stream, &graph), "end capture") ||
          hip_check(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "instantiate"))
        rc = GFX_ERR_DEVICE;
    }
    while (rc == GFX_OK && done + steps_per_graph <= nsteps) {
      if (hip_check(hipGraphLaunch(exec, s.stream), "graph launch")) rc = GFX_ERR_DEVICE;
      done += steps_per_graph;
      b.graph_launches++;
    }
    if (exec) hipGraphExecDestroy(exec);
    if (graph) hipGraphDestroy(graph);
  }
  while (rc == GFX_OK && done < nsteps) {
    one_step();
    done++;
  }
  if (hip_check(hipStreamSynchronize(s.stream), "anneal") && rc == GFX_OK) rc = GFX_ERR_DEVICE;
  b.run_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  hipFree(d_betas);
  hipFree(d_step);
  return rc;
}

// the lattice with the largest F among those sharing lattice base's links
// (after gfx_batch_run): the Gribov copy to select
int gfx_batch_best_copy(int base, int* q_best, double* F_best) {
  if (base < 0 || base >= (int)g_b.probs.size() || !g_b.h_results) return GFX_ERR_STATE;
  int best = -1;
  double fb = -1e300;
  for (int q = 0; q < (int)g_b.probs.size(); q++)
    if (g_b.probs[q].ushare == g_b.probs[base].ushare && g_b.h_results[4 * q] > fb) {
      fb = g_b.h_results[4 * q];
      best = q;
    }
  if (q_best) *q_best = best;
  if (F_best) *F_best = fb;
  return GFX_OK;
}

// F, dAdA, max dAdA and the stop iteration (-1: not converged) of lattice q
int gfx_batch_result(int q, double* F, double* dAdA, double* maxdAdA, int* iter) {
  if (q < 0 || q >= (int)g_b.probs.size() || !g_b.h_results) return GFX_ERR_STATE;
  if (F) *F = g_b.h_results[4 * q];
  if (dAdA) *dAdA = g_b.h_results[4 * q + 1];
  if (maxdAdA) *maxdAdA = g_b.h_results[4 * q + 2];
  if (iter) *iter = g_b.probs[q].stop_iter[1];
  return GFX_OK;
}

int gfx_batch_get_trafo(int q, gfx_su3* trafo) {
  if (q < 0 || q >= (int)g_b.probs.size()) return GFX_ERR_STATE;
  return download_aos_geo(g_st, g_b.probs[q].geo, g_b.probs[q].g64, (double*)trafo, 1);
}

// wall time of the last gfx_batch_run, graph launches, iterations in each
// precision and the site updates done (sum over lattices of V x iterations
// while the lattice was being updated)
int gfx_batch_stats(double* ms, int* graph_launches, int* iters32, int* iters64,
                    double* site_updates) {
  if (ms) *ms = g_b.run_ms;
  if (graph_launches) *graph_launches = g_b.graph_launches;
  if (iters32) *iters32 = g_b.iters32;
  if (iters64) *iters64 = g_b.iters64;
  if (site_updates) *site_updates = g_b.site_updates;
  return GFX_OK;
}

// the sweep variant picked for the batch in each precision (index 0 FP64, 1 FP32)
int gfx_batch_variant(int precision, int* variant, double* ms_per_variant) {
  const int pi = precision == GFX_PREC_MIXED ? 1 : 0;
  if (variant) *variant = g_b.variant[pi];
  if (ms_per_variant)
    for (int k = 0; k < SWEEP_VARIANTS; k++) ms_per_variant[k] = g_b.tune_ms[pi][k];
  return g_b.tuned[pi] ? GFX_OK : GFX_ERR_STATE;
}

}  // extern "C"
