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
  PackedArgs pk{};  // packed links (owner of the links only; mixed runs)
  double2* g64 = nullptr;
  float4* g32 = nullptr;
  // 32-bit fixed-point links (owner only; 8 fields) and g (3 fields) of the
  // FP64 phase of mixed runs
  int4* Ufx = nullptr;
  int4* gfx = nullptr;
  int ushare = 0;  // problem whose links this one uses (itself if it owns them)
  int stop_iter = -1;  // iteration of the stop check (-1: not converged)
  int switch_iter = -1;  // mixed: iteration of the check that left FP32
  int idle = 0;  // iterations it waited for the FP64 sweeps to be launched
};

// Index 0 = FP64, 1 = FP32 for every per-precision array.
struct Batch {
  std::vector<BProbHost> probs;
  BatchProblem* d_probs[2] = {nullptr, nullptr};  // field pointers in each precision
  int2* d_bmap[2][TILINGS] = {};  // per [block 256/64][tiling], every lattice
  int nbmap[2] = {0, 0};
  std::vector<int2> h_order[2][TILINGS];  // the same lists before dealing to XCDs
  // per [precision][block][tiling]: the lattices sweeping in that precision
  // only, rebuilt when the census version moves on
  struct CMap {
    int2* d = nullptr;
    int n = 0;
    long version = -1;
    std::vector<int2> h;
  } cmap[2][2][TILINGS];
  long census_version = 0;
  int* d_poff[2] = {nullptr, nullptr};  // check-partial offsets per problem, per block size
  double* d_pod[2] = {nullptr, nullptr};  // check partials of the odd sweeps
  double* d_pev[2] = {nullptr, nullptr};  // and of the even sweeps
  int* d_act[2] = {nullptr, nullptr};     // lattice is sweeping in this precision
  int* d_inmap[2] = {nullptr, nullptr};   // lattice is in the precision's maps (last census)
  int* h_act = nullptr;                   // pinned, 2 x n
  std::vector<int> h_act_prev;            // the census before
  int* d_conv = nullptr;                  // FP32 -> FP64 conversions of a check
  int* d_sw = nullptr;                    // iteration of each lattice's switch
  int* h_sw = nullptr;
  int* d_roff = nullptr;      // per problem: first site of its random numbers
  int nh_total = 0;           // sites per parity over all problems
  double* d_best = nullptr;   // per problem: best max dAdA of the FP32 phase
  int* d_stall = nullptr;     // checks since it last halved
  double* d_results = nullptr;
  double* h_results = nullptr;
  double* d_fres = nullptr;   // functionals of the final states (4 per problem)
  int* d_iter = nullptr;
  bool built = false;
  int variant[2] = {0, 0};
  bool tuned[2] = {false, false};
  bool packed = false;         // the current run sweeps on packed links
  bool fixed = false;          // and its FP64 phase on fixed-point links and g
  bool anneal_fp32 = false;    // gfx_batch_anneal on FP32 links and g
  int tuned_links = -1;        // the link form (BLINK_*) the variants were timed on
  float tune_ms[2][SWEEP_VARIANTS] = {};
  // check periods in persistent launches (gfx_batch_set_persistent) where
  // every lattice still sweeping is in one precision
  bool persistent = false;
  int* d_changed = nullptr;  // a lattice stopped or left FP32 in the launch
  int* d_done = nullptr;     // check periods the launch ran
  int* h_done = nullptr;
  // statistics of the last run
  double run_ms = 0.0;
  int graph_launches = 0, iters32 = 0, iters64 = 0;
  int persistent_launches = 0;
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
      b.h_order[k][t].clear();
      for (int pr = 0; pr < 2; pr++) {
        hipFree(b.cmap[pr][k][t].d);
        b.cmap[pr][k][t].d = nullptr;
        b.cmap[pr][k][t].n = 0;
        b.cmap[pr][k][t].version = -1;
      }
    }
    hipFree(b.d_poff[k]);
    hipFree(b.d_probs[k]);
    hipFree(b.d_pod[k]);
    hipFree(b.d_pev[k]);
    hipFree(b.d_act[k]);
    hipFree(b.d_inmap[k]);
    b.d_inmap[k] = nullptr;
    b.d_poff[k] = nullptr;
    b.d_probs[k] = nullptr;
    b.d_pod[k] = b.d_pev[k] = nullptr;
    b.d_act[k] = nullptr;
  }
  hipFree(b.d_conv);
  hipFree(b.d_sw);
  hipFree(b.d_best);
  hipFree(b.d_roff);
  hipFree(b.d_stall);
  hipFree(b.d_results);
  hipFree(b.d_fres);
  hipFree(b.d_iter);
  hipFree(b.d_changed);
  hipFree(b.d_done);
  if (b.h_done) hipHostFree(b.h_done);
  b.d_changed = b.d_done = b.h_done = nullptr;
  if (b.h_act) hipHostFree(b.h_act);
  if (b.h_sw) hipHostFree(b.h_sw);
  if (b.h_results) hipHostFree(b.h_results);
  b.d_conv = b.d_sw = nullptr;
  b.d_best = nullptr;
  b.d_roff = nullptr;
  b.d_stall = nullptr;
  b.d_results = nullptr;
  b.d_fres = nullptr;
  b.d_iter = nullptr;
  b.h_act = b.h_sw = nullptr;
  b.h_results = nullptr;
  b.built = false;
}

static void batch_free(Batch& b) {
  if (g_st.stream) hipStreamSynchronize(g_st.stream);
  for (auto& p : b.probs) {
    hipFree(p.g64);
    hipFree(p.g32);
    hipFree(p.gfx);
    if (&p - &b.probs[0] == p.ushare) {
      hipFree(p.U64);
      hipFree(p.U32);
      hipFree(p.pk.a);
      hipFree(p.Ufx);
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
      if (hip_check(hipMemcpy(b.d_bmap[k][t], map.data(), sizeof(int2) * map.size(), hipMemcpyHostToDevice), "batch map"))
        return GFX_ERR_DEVICE;
      b.h_order[k][t] = list;
      for (int pr = 0; pr < 2; pr++)
        if (dev_alloc((void**)&b.cmap[pr][k][t].d, sizeof(int2) * map.size(), "batch map")) return GFX_ERR_ALLOC;
    }
    std::vector<int> poff(n + 1, 0);
    for (int q = 0; q < n; q++) poff[q + 1] = poff[q] + (b.probs[q].geo.nh + BLOCKS[k] - 1) / BLOCKS[k];
    if (dev_alloc((void**)&b.d_poff[k], sizeof(int) * (n + 1), "batch offsets")) return GFX_ERR_ALLOC;
    if (hip_check(hipMemcpy(b.d_poff[k], poff.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice), "batch offsets"))
      return GFX_ERR_DEVICE;
  }
  std::vector<int> roff(n + 1, 0);
  for (int q = 0; q < n; q++) roff[q + 1] = roff[q] + b.probs[q].geo.nh;
  b.nh_total = roff[n];
  if (dev_alloc((void**)&b.d_roff, sizeof(int) * (n + 1), "batch random offsets")) return GFX_ERR_ALLOC;
  if (hip_check(hipMemcpy(b.d_roff, roff.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice), "batch offsets"))
    return GFX_ERR_DEVICE;
  // partial slots: as many as blocks of the smallest workgroup
  size_t npart = 0;
  for (int q = 0; q < n; q++) npart += (b.probs[q].geo.nh + 63) / 64;
  for (int k = 0; k < 2; k++)
    if (dev_alloc((void**)&b.d_pod[k], sizeof(double) * 3 * npart, "batch partials") ||
        dev_alloc((void**)&b.d_pev[k], sizeof(double) * 3 * npart, "batch partials") ||
        dev_alloc((void**)&b.d_probs[k], sizeof(BatchProblem) * n, "batch problems") ||
        dev_alloc((void**)&b.d_act[k], sizeof(int) * n, "batch flags") ||
        dev_alloc((void**)&b.d_inmap[k], sizeof(int) * n, "batch flags"))
      return GFX_ERR_ALLOC;
  if (dev_alloc((void**)&b.d_best, sizeof(double) * n, "batch best") ||
      dev_alloc((void**)&b.d_stall, sizeof(int) * n, "batch stall") ||
      dev_alloc((void**)&b.d_conv, sizeof(int) * n, "batch conversions") ||
      dev_alloc((void**)&b.d_sw, sizeof(int) * n, "batch switches") ||
      dev_alloc((void**)&b.d_results, sizeof(double) * 4 * n, "batch results") ||
      dev_alloc((void**)&b.d_fres, sizeof(double) * 4 * n, "batch results") ||
      dev_alloc((void**)&b.d_iter, sizeof(int), "batch iter") ||
      dev_alloc((void**)&b.d_changed, sizeof(int), "batch flags") ||
      dev_alloc((void**)&b.d_done, sizeof(int), "batch flags"))
    return GFX_ERR_ALLOC;
  if (hipHostMalloc((void**)&b.h_act, sizeof(int) * 2 * n, 0) != hipSuccess ||
      hipHostMalloc((void**)&b.h_sw, sizeof(int) * n, 0) != hipSuccess ||
      hipHostMalloc((void**)&b.h_results, sizeof(double) * 4 * n, 0) != hipSuccess ||
      hipHostMalloc((void**)&b.h_done, sizeof(int), 0) != hipSuccess)
    return GFX_ERR_ALLOC;
  b.built = true;
  return GFX_OK;
}

template <typename V>
static V* bfield(V* g, const Geom& geo, int k) {
  return g + (size_t)k * Store<V>::NPLANE * geo.stride;
}

// canonical field pointers in both precisions: even = field 0, odd =
// field 1, spare = field 2 (FP32 pointers only where FP32 fields exist; the
// FP64 phase's fields are the fixed-point ones in a run with b.fixed)
static void set_pointers(Batch& b) {
  const size_t n = b.probs.size();
  std::vector<BatchProblem> h64(n), h32(n);
  for (size_t q = 0; q < n; q++) {
    const BProbHost& p = b.probs[q];
    const BProbHost& o = b.probs[p.ushare];
    h64[q].geo = h32[q].geo = p.geo;
    h64[q].U = b.fixed ? (const void*)o.Ufx : (const void*)o.U64;
    h32[q].U = o.U32;
    const PackedLinks pk = b.packed ? PackedLinks{o.pk.a, o.pk.b, o.pk.c, o.pk.lo, o.pk.a21, o.pk.b21} : PackedLinks{};
    h64[q].pk = b.fixed ? PackedLinks{} : pk;
    h32[q].pk = pk;
    for (int k = 0; k < 3; k++) {
      void* f64 = b.fixed ? (void*)bfield(p.gfx, p.geo, k) : (void*)bfield(p.g64, p.geo, k);
      void* f32 = p.g32 ? (void*)bfield(p.g32, p.geo, k) : nullptr;
      if (k < 2) {
        h64[q].g[k] = f64;
        h32[q].g[k] = f32;
      } else {
        h64[q].galt = f64;
        h32[q].galt = f32;
      }
    }
  }
  hipStreamSynchronize(g_st.stream);
  hipMemcpy(b.d_probs[0], h64.data(), sizeof(BatchProblem) * n, hipMemcpyHostToDevice);
  hipMemcpy(b.d_probs[1], h32.data(), sizeof(BatchProblem) * n, hipMemcpyHostToDevice);
}

// after a run: copy each FP64 even field the checks left in the spare slot
// back to field 0.  Fixed-point g (b.fixed) then goes to the FP64 fields,
// reunitarised (the state the stop bound holds for), and each lattice's F,
// dAdA and max dAdA are evaluated again on its FP64 links and g (the checks
// saw the rounded fields); the stop iterations stay.
static void batch_canonical(Batch& b) {
  State& s = g_st;
  std::vector<BatchProblem> h(b.probs.size());
  hipStreamSynchronize(s.stream);
  hipMemcpy(h.data(), b.d_probs[0], sizeof(BatchProblem) * h.size(), hipMemcpyDeviceToHost);
  for (size_t q = 0; q < b.probs.size(); q++) {
    const BProbHost& p = b.probs[q];
    void* f0 = b.fixed ? (void*)p.gfx : (void*)p.g64;
    const size_t bytes = b.fixed ? field_bytes_fx(p.geo) : field_bytes64(p.geo);
    if (h[q].g[0] != f0) hipMemcpyAsync(f0, h[q].g[0], bytes, hipMemcpyDeviceToDevice, s.stream);
    if (!b.fixed) continue;
    launch_fx_to_g64(p.gfx, p.g64, p.geo, s.stream);
    // partials: the FP64 odd-check slots, ceil(nh / 64) blocks per lattice
    // in all, at least the functional's blocks of the largest lattice
    launch_functional_f64(b.probs[p.ushare].U64, bfield(p.g64, p.geo, 0), bfield(p.g64, p.geo, 1),
                          p.geo, 4, 1, b.d_pod[0], 0.0, b.d_fres + 4 * q, s.stream);
    hipMemcpyAsync(b.d_results + 4 * q, b.d_fres + 4 * q, 3 * sizeof(double),
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

// the packed links of every link set (as ensure_packed_links, 60 bytes per
// link in one allocation), made from the FP64 links
static int ensure_packed(Batch& b) {
  State& s = g_st;
  for (size_t q = 0; q < b.probs.size(); q++) {
    BProbHost& p = b.probs[q];
    if ((int)q != p.ushare) continue;
    if (!p.pk.a) {
      const size_t n = (size_t)8 * p.geo.stride;
      char* base = nullptr;
      if (dev_alloc((void**)&base, n * (16 + 16 + 4 + 24 + 32), "batch packed links")) return GFX_ERR_ALLOC;
      p.pk.a = reinterpret_cast<uint4*>(base);
      p.pk.b = reinterpret_cast<uint4*>(base + n * 16);
      p.pk.c = reinterpret_cast<uint32_t*>(base + n * 32);
      p.pk.lo = reinterpret_cast<uint2*>(base + n * 36);
      p.pk.a21 = reinterpret_cast<uint4*>(base + n * 60);
      p.pk.b21 = reinterpret_cast<uint4*>(base + n * 76);
    }
    launch_pack_links(p.U64, p.pk, p.geo, 8, s.stream);
  }
  return GFX_OK;
}

// the fixed-point links of every link set (8 fields) and the fixed-point
// g fields (3) of every lattice, made from the FP64 links; g is written by
// the conversions out of FP32
static int ensure_fixed(Batch& b) {
  State& s = g_st;
  for (size_t q = 0; q < b.probs.size(); q++) {
    BProbHost& p = b.probs[q];
    if ((int)q == p.ushare) {
      if (!p.Ufx && dev_alloc((void**)&p.Ufx, 8 * field_bytes_fx(p.geo), "batch fixed-point links"))
        return GFX_ERR_ALLOC;
      launch_f64_to_fx(p.U64, p.Ufx, p.geo, 8, s.stream);
    }
    if (!p.gfx && dev_alloc((void**)&p.gfx, 3 * field_bytes_fx(p.geo), "batch fixed-point trafo"))
      return GFX_ERR_ALLOC;
  }
  return GFX_OK;
}

// the workgroup map of precision pr for variant v: the blocks of the
// lattices sweeping in pr at the last census, in the variant's order, dealt
// to the XCDs (stopped or other-precision lattices cost no workgroups)
static const Batch::CMap& batch_map(Batch& b, int pr, int v) {
  State& s = g_st;
  const int k = vblock_index(v), t = sweep_variant_tiling(v);
  Batch::CMap& m = b.cmap[pr][k][t];
  if (m.version != b.census_version) {
    const size_t n = b.probs.size();
    std::vector<int2> list;
    list.reserve(b.h_order[k][t].size());
    for (const int2& e : b.h_order[k][t])
      if (b.h_act[pr * n + e.x]) list.push_back(e);
    deal_to_xcds(list, m.h);
    m.n = (int)m.h.size();
    if (m.n > 0) hipMemcpyAsync(m.d, m.h.data(), sizeof(int2) * m.n, hipMemcpyHostToDevice, s.stream);
    m.version = b.census_version;
  }
  return m;
}

static int links_of(const Batch& b) {
  return b.fixed ? BLINK_FIXED : b.packed ? BLINK_PACKED : BLINK_F64;
}

// one half sweep in precision pr (0 FP64, 1 FP32) of the lattices sweeping
// in it; chk as in launch_batch_sweep
static void batch_half(Batch& b, int pr, int parity, int chk) {
  State& s = g_st;
  const int v = b.variant[pr];
  const Batch::CMap& m = batch_map(b, pr, v);
  if (m.n == 0) return;
  launch_batch_sweep(pr == 1, b.d_probs[pr], m.d, m.n, b.d_act[pr], b.d_poff[vblock_index(v)],
                     parity, s.omega, v, chk, chk == 2 ? b.d_pev[pr] : b.d_pod[pr], s.stream,
                     links_of(b));
}

// check periods per persistent launch at most (the host sees the census
// between launches)
static const int PERSIST_PERIODS = 32;

struct RunParams {
  double eps, switch_theta;
  int ck, maxit, periods, stall_checks;
};

// the end of the check of iteration *d_iter: the even sweeps of the next
// iteration carrying the check, the per-lattice grid pass, and the
// conversions of the lattices that leave FP32
static BatchCheck batch_check_args(Batch& b) {
  BatchCheck c;
  for (int pr = 0; pr < 2; pr++) {
    c.pod[pr] = b.d_pod[pr];
    c.pev[pr] = b.d_pev[pr];
    c.poff[pr] = b.d_poff[vblock_index(b.variant[pr])];
    c.probs[pr] = b.d_probs[pr];
    c.act[pr] = b.d_act[pr];
    c.inmap[pr] = b.d_inmap[pr];
  }
  c.conv = b.d_conv;
  c.sw = b.d_sw;
  c.results = b.d_results;
  c.iter = b.d_iter;
  c.best = b.d_best;
  c.stall = b.d_stall;
  c.links = links_of(b);
  return c;
}

// the conversions of the lattices that left FP32 at the last check
static void batch_convert_left(Batch& b) {
  State& s = g_st;
  const int k = block_index(batch_convert_block());
  launch_batch_convert(b.d_probs[1], b.d_probs[0], b.d_bmap[k][0], b.nbmap[k], b.d_conv, s.stream,
                       b.fixed);
  launch_fill_i32(b.d_conv, (int)b.probs.size(), 0, s.stream);
}

static void batch_check_tail(Batch& b, const bool use[2], const RunParams& rp) {
  State& s = g_st;
  for (int pr = 0; pr < 2; pr++)
    if (use[pr]) batch_half(b, pr, 0, 2);
  const BatchCheck c = batch_check_args(b);
  const int n = (int)b.probs.size();
  launch_batch_finalize(c, n, 4, rp.eps, rp.switch_theta, rp.maxit, rp.stall_checks, s.stall_gate,
                        s.stream);
  if (use[1]) batch_convert_left(b);
}

// lattices sweeping in each precision (reads the device flags)
static int census(Batch& b, int cnt[2]) {
  State& s = g_st;
  const size_t n = b.probs.size();
  hipMemcpyAsync(b.h_act, b.d_act[0], sizeof(int) * n, hipMemcpyDeviceToHost, s.stream);
  hipMemcpyAsync(b.h_act + n, b.d_act[1], sizeof(int) * n, hipMemcpyDeviceToHost, s.stream);
  hipMemcpyAsync(b.h_sw, b.d_sw, sizeof(int) * n, hipMemcpyDeviceToHost, s.stream);
  if (hip_check(hipStreamSynchronize(s.stream), "batch flags")) return GFX_ERR_DEVICE;
  cnt[0] = cnt[1] = 0;
  bool changed = false;
  for (size_t q = 0; q < n; q++) {
    cnt[0] += b.h_act[q] ? 1 : 0;
    cnt[1] += b.h_act[n + q] ? 1 : 0;
    changed |= (b.h_act[q] != b.h_act_prev[q]) || (b.h_act[n + q] != b.h_act_prev[n + q]);
    b.h_act_prev[q] = b.h_act[q];
    b.h_act_prev[n + q] = b.h_act[n + q];
  }
  if (changed) {
    b.census_version++;
    // the maps rebuilt from this census hold exactly these lattices
    for (int pr = 0; pr < 2; pr++)
      hipMemcpyAsync(b.d_inmap[pr], b.h_act + pr * n, sizeof(int) * n, hipMemcpyHostToDevice, s.stream);
  }
  return GFX_OK;
}

// The run: every lattice sweeps in FP32 (mixed) or FP64 from iteration 0;
// a check every ck iterations may stop it or move it from FP32 to FP64.
// Eager iterations (which also time the sweep variants of a precision the
// first time it has lattices) run up to the iteration after a check; then a
// graph of `periods` check periods, with the sweeps of the precisions that
// have lattices, is captured and replayed while it ends before maxit and no
// lattice changed precision; then back to eager iterations.  A lattice that
// leaves FP32 inside a replay whose graph has no FP64 sweeps waits for the
// replay to end (its `idle` count; its stop iteration is reported without
// the wait).
static int batch_loop(Batch& b, bool mixed, const RunParams& rp, int* it_end) {
  State& s = g_st;
  const int n = (int)b.probs.size();
  const int ck = rp.ck;
  set_pointers(b);
  launch_fill_i32(b.d_act[0], n, mixed ? 0 : 1, s.stream);
  launch_fill_i32(b.d_act[1], n, mixed ? 1 : 0, s.stream);
  launch_fill_i32(b.d_conv, n, 0, s.stream);
  launch_fill_i32(b.d_sw, n, -1, s.stream);
  launch_fill_i32(b.d_stall, n, 0, s.stream);
  launch_fill_f64(b.d_best, n, 1e300, s.stream);
  launch_fill_f64(b.d_results, 4 * n, -1.0, s.stream);
  int cnt[2] = {mixed ? 0 : n, mixed ? n : 0};
  b.h_act_prev.assign(2 * n, 0);
  for (int q = 0; q < n; q++) {
    b.h_act[q] = b.h_act_prev[q] = mixed ? 0 : 1;
    b.h_act[n + q] = b.h_act_prev[n + q] = mixed ? 1 : 0;
  }
  b.census_version++;
  for (int pr = 0; pr < 2; pr++)
    hipMemcpyAsync(b.d_inmap[pr], b.h_act + pr * n, sizeof(int) * n, hipMemcpyHostToDevice, s.stream);
  int i = 0, rc = GFX_OK;
  bool even_done = false;  // the even half of iteration i ran with the last check
  int tune_v[2] = {b.tuned[0] ? -1 : 0, b.tuned[1] ? -1 : 0};
  std::vector<int> sw_before(n, -1);
  // after a check (census just read): a lattice that left FP32 inside a
  // graph replay (whose FP64 map did not hold it) waited until `resume`
  auto after_check = [&](int resume, bool in_graph) {
    for (int q = 0; q < n; q++) {
      const int sw = b.h_sw[q];
      if (sw >= 0 && sw_before[q] < 0 && in_graph && b.h_act[q] && resume > sw + 1)
        b.probs[q].idle += resume - (sw + 1);
      sw_before[q] = sw;
    }
  };
  // one eager iteration
  auto eager = [&]() {
    const bool chk = (i % ck) == 0;
    const bool both = !even_done;
    const bool use[2] = {cnt[0] > 0, cnt[1] > 0};
    int keep[2] = {b.variant[0], b.variant[1]};
    bool timed[2] = {false, false};
    for (int pr = 0; pr < 2; pr++) {
      if (!use[pr]) continue;
      // time full non-check iterations when there are any (ck >= 3), else
      // every iteration (the check sweeps then use the variant being timed)
      timed[pr] = tune_v[pr] >= 0 && ((ck >= 3) ? (both && !chk) : true);
      if (timed[pr]) b.variant[pr] = tune_v[pr];
      if (timed[pr]) hipEventRecord(s.ev0, s.stream);
      if (both) batch_half(b, pr, 0, 0);
      batch_half(b, pr, 1, chk ? 1 : 0);
      if (timed[pr]) {
        hipEventRecord(s.ev1, s.stream);
        hipEventSynchronize(s.ev1);
        float t = 0.f;
        hipEventElapsedTime(&t, s.ev0, s.ev1);
        b.tune_ms[pr][tune_v[pr]] = both ? 0.5f * t : t;
      }
      if (pr == 1) b.iters32++; else b.iters64++;
    }
    even_done = false;
    if (chk) {
      launch_fill_i32(b.d_iter, 1, i, s.stream);
      batch_check_tail(b, use, rp);
      even_done = true;
      if (census(b, cnt)) rc = GFX_ERR_DEVICE;
      after_check(i + 1, false);
    }
    for (int pr = 0; pr < 2; pr++) {
      if (!timed[pr]) continue;
      b.variant[pr] = keep[pr];
      if (++tune_v[pr] == SWEEP_VARIANTS) {
        int best = 0;
        for (int k = 1; k < SWEEP_VARIANTS; k++)
          if (b.tune_ms[pr][k] < b.tune_ms[pr][best]) best = k;
        b.variant[pr] = best;
        b.tuned[pr] = true;
        tune_v[pr] = -1;
      }
    }
    i++;
  };
  // persistent launches of the check periods while every lattice still
  // sweeping is in precision pr: each runs until a lattice stops or leaves
  // FP32 (its conversion follows the launch) or PERSIST_PERIODS periods
  auto persistent_stretch = [&](int pr) {
    const Batch::CMap& m = batch_map(b, pr, b.variant[pr]);
    const long version0 = b.census_version;
    launch_fill_i32(b.d_iter, 1, i + ck - 1, s.stream);
    while (rc == GFX_OK && i + ck <= rp.maxit) {
      PersistArgs a{};
      a.bmap = m.d;
      a.nblocks = m.n;
      a.pod = b.d_pod[pr];
      a.pev = b.d_pev[pr];
      a.nprob = n;
      a.ck = ck;
      a.periods = std::min(PERSIST_PERIODS, (rp.maxit - i) / ck);
      a.eps = rp.eps;
      a.switch_theta = rp.switch_theta;
      a.maxit = rp.maxit;
      a.stall_checks = rp.stall_checks;
      a.stall_gate = s.stall_gate;
      a.changed = b.d_changed;
      a.done = b.d_done;
      launch_fill_i32(b.d_changed, 1, 0, s.stream);
      launch_fill_i32(b.d_done, 1, 0, s.stream);
      if (hip_check(launch_batch_persistent(pr == 1, links_of(b), b.variant[pr], a,
                                            batch_check_args(b), s.omega, s.stream),
                    "persistent launch")) {
        rc = GFX_ERR_DEVICE;
        break;
      }
      b.persistent_launches++;
      hipMemcpyAsync(b.h_done, b.d_done, sizeof(int), hipMemcpyDeviceToHost, s.stream);
      if (pr == 1) batch_convert_left(b);
      int c2[2];
      if (census(b, c2)) {
        rc = GFX_ERR_DEVICE;
        break;
      }
      const int done = *b.h_done;
      if (done < 1) {  // a launch runs one period at least
        rc = GFX_ERR_DEVICE;
        break;
      }
      i += done * ck;
      if (pr == 1) b.iters32 += done * ck; else b.iters64 += done * ck;
      after_check(i, true);
      cnt[0] = c2[0];
      cnt[1] = c2[1];
      if (cnt[0] + cnt[1] == 0 || b.census_version != version0) break;
    }
  };
  while (rc == GFX_OK && i < rp.maxit && cnt[0] + cnt[1] > 0) {
    // eager until tuned (for the precisions in use) and just after a check
    do {
      eager();
    } while (rc == GFX_OK && i < rp.maxit && cnt[0] + cnt[1] > 0 &&
             !(even_done && (cnt[0] == 0 || tune_v[0] < 0) && (cnt[1] == 0 || tune_v[1] < 0)));
    if (rc != GFX_OK || cnt[0] + cnt[1] == 0) continue;
    if (b.persistent && (cnt[0] > 0) != (cnt[1] > 0) && i + ck <= rp.maxit) {
      persistent_stretch(cnt[0] > 0 ? 0 : 1);
      continue;
    }
    if (i + rp.periods * ck > rp.maxit) continue;
    // graph of `periods` check periods for the precisions in use
    const bool use[2] = {cnt[0] > 0, cnt[1] > 0};
    // maps are uploaded before the capture (no copy may enter the graph)
    for (int pr = 0; pr < 2; pr++)
      if (use[pr]) batch_map(b, pr, b.variant[pr]);
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    if (hip_check(hipStreamBeginCapture(s.stream, hipStreamCaptureModeThreadLocal), "capture")) return GFX_ERR_DEVICE;
    for (int per = 0; per < rp.periods; per++) {
      for (int j = 0; j < ck; j++) {
        for (int pr = 0; pr < 2; pr++)
          if (use[pr] && j > 0) batch_half(b, pr, 0, 0);
        for (int pr = 0; pr < 2; pr++)
          if (use[pr]) batch_half(b, pr, 1, j == ck - 1 ? 1 : 0);
      }
      batch_check_tail(b, use, rp);
      launch_batch_advance(b.d_iter, ck, s.stream);
    }
    if (hip_check(hipStreamEndCapture(s.stream, &graph), "end capture")) return GFX_ERR_DEVICE;
    if (hip_check(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "instantiate")) {
      hipGraphDestroy(graph);
      return GFX_ERR_DEVICE;
    }
    launch_fill_i32(b.d_iter, 1, i + ck - 1, s.stream);
    const long version_at_capture = b.census_version;
    while (rc == GFX_OK && i + rp.periods * ck <= rp.maxit) {
      if (hip_check(hipGraphLaunch(exec, s.stream), "graph launch")) {
        rc = GFX_ERR_DEVICE;
        break;
      }
      b.graph_launches++;
      i += rp.periods * ck;
      if (use[1]) b.iters32 += rp.periods * ck;
      if (use[0]) b.iters64 += rp.periods * ck;
      int c2[2];
      if (census(b, c2)) {
        rc = GFX_ERR_DEVICE;
        break;
      }
      after_check(i, true);
      cnt[0] = c2[0];
      cnt[1] = c2[1];
      if (cnt[0] + cnt[1] == 0) break;
      if (b.census_version != version_at_capture) break;  // a lattice stopped or switched
    }
    hipGraphExecDestroy(exec);
    hipGraphDestroy(graph);
  }
  if (mixed && rc == GFX_OK) {
    // lattices still in FP32 at maxit: their checked state to FP64
    int c2[2];
    if (census(b, c2) == GFX_OK && c2[1] > 0) {
      std::vector<int> cv(n);
      for (int q = 0; q < n; q++) cv[q] = b.h_act[n + q] ? 2 : 0;
      hipMemcpy(b.d_conv, cv.data(), sizeof(int) * n, hipMemcpyHostToDevice);
      const int k = block_index(batch_convert_block());
      launch_batch_convert(b.d_probs[1], b.d_probs[0], b.d_bmap[k][0], b.nbmap[k], b.d_conv,
                           s.stream, b.fixed);
      launch_fill_i32(b.d_conv, n, 0, s.stream);
    }
  }
  batch_canonical(b);
  *it_end = i;
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
  if (!geometry_ok(l0, l1, l2, l3)) return -GFX_ERR_GEOMETRY;
  if (ensure_common(0)) return -GFX_ERR_DEVICE;
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
  b.packed = mixed && s.packed_links && eps > PACK_EPS_MIN;
  if (b.packed && ensure_packed(b)) return GFX_ERR_ALLOC;
  // the FP64 phase of a run on packed links sweeps on fixed-point links and
  // g when eps is well above the fixed-point bound its stops carry
  b.fixed = b.packed && eps > FX_EPS_BOUND_MIN;
  if (b.fixed && ensure_fixed(b)) return GFX_ERR_ALLOC;
  if (links_of(b) != b.tuned_links) {
    // the sweep variants are timed again on the other link form
    b.tuned[0] = b.tuned[1] = false;
    b.tuned_links = links_of(b);
  }
  b.graph_launches = b.iters32 = b.iters64 = b.persistent_launches = 0;
  for (auto& p : b.probs) {
    p.stop_iter = p.switch_iter = -1;
    p.idle = 0;
  }
  RunParams rp;
  rp.eps = eps;
  rp.switch_theta = mixed ? std::max(switch_theta > 0 ? switch_theta : s.switch_theta, eps) : 0.0;
  rp.ck = checkinterval;
  rp.maxit = maxit;
  rp.periods = periods_per_graph;
  rp.stall_checks = mixed ? (s.stall_iters + checkinterval - 1) / checkinterval : 0;
  hipStreamSynchronize(s.stream);
  const auto t0 = std::chrono::steady_clock::now();
  int it = 0;
  int rc = batch_loop(b, mixed, rp, &it);
  hipMemcpyAsync(b.h_results, b.d_results, sizeof(double) * 4 * b.probs.size(), hipMemcpyDeviceToHost, s.stream);
  hipMemcpyAsync(b.h_sw, b.d_sw, sizeof(int) * b.probs.size(), hipMemcpyDeviceToHost, s.stream);
  hipStreamSynchronize(s.stream);
  b.run_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  int nconv = 0;
  b.site_updates = 0.0;
  for (size_t q = 0; q < b.probs.size(); q++) {
    BProbHost& p = b.probs[q];
    const int stop = (int)b.h_results[4 * q + 3];
    p.switch_iter = b.h_sw[q];
    // iterations this lattice ran: up to its stop check, without waits
    p.stop_iter = stop >= 0 ? stop - p.idle : -1;
    if (stop >= 0) nconv++;
    const int ran = (stop >= 0 ? stop + 1 : it) - p.idle;
    b.site_updates += 2.0 * p.geo.nh * (double)ran;
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
  // the anneal works on the FP64 (or FP32) fields, whatever the last run's
  // FP64 phase swept on
  b.fixed = false;
  set_pointers(b);
  std::vector<double> betas(nsteps > 0 ? nsteps : 1);
  for (int i = 0; i < nsteps; i++) betas[i] = 1.0 / gfx_sa_temperature(i, Tmin, Tmax, expo, nsteps);
  double* d_betas = nullptr;
  int* d_step = nullptr;
  if (dev_alloc((void**)&d_betas, sizeof(double) * betas.size(), "SA schedule") ||
      dev_alloc((void**)&d_step, sizeof(int), "SA step")) {
    hipFree(d_betas);
    return GFX_ERR_ALLOC;
  }
  if (hip_check(hipMemcpy(d_betas, betas.data(), sizeof(double) * betas.size(), hipMemcpyHostToDevice), "SA schedule")) {
    hipFree(d_betas);
    hipFree(d_step);
    return GFX_ERR_DEVICE;
  }
  launch_fill_i32(d_step, 1, 0, s.stream);
  const int k = block_index(sa_block_size());
  const int2* map = b.d_bmap[k][0];
  const int nblk = b.nbmap[k];
  // FP32 anneal: every lattice's g (and its links) in FP32 for the whole
  // schedule, reunitarised into FP64 at the end
  const bool f32 = b.anneal_fp32;
  if (f32) {
    if ((rc = ensure_fp32(b))) {
      hipFree(d_betas);
      hipFree(d_step);
      return rc;
    }
    set_pointers(b);
  }
  const BatchProblem* probs = b.d_probs[f32 ? 1 : 0];
  if ((rc = mt_second_set(s))) {
    hipFree(d_betas);
    hipFree(d_step);
    return rc;
  }
  // The SA half sweep of parity p reads fields[p]; the numbers of the next
  // SA half sweep are made into the other fields on the twister stream while
  // this one (and, after the odd one, the microcanonical sweeps, which draw
  // no numbers) runs, once the half sweep before has read them.  Refills stay
  // in order on one stream, so every half sweep gets the numbers it got with
  // a refill after each.  A stretch of steps starts with fields[0] holding
  // the current numbers and ends with the twister stream joined back (an
  // even number of half sweeps leaves the current numbers in fields[0]).
  float* fg[2] = {s.rnd_gauss, s.rnd_gauss2};
  float* fu[2] = {s.rnd_unif, s.rnd_unif2};
  auto fork = [&]() {
    hipEventRecord(s.ev_filled[0], s.stream);
    hipStreamWaitEvent(s.mt_stream, s.ev_filled[0], 0);
  };
  auto join = [&]() { hipStreamWaitEvent(s.stream, s.ev_filled[0], 0); };
  auto one_step = [&](bool first) {
    for (int p = 0; p < 2; p++) {
      const bool start = first && p == 0;
      if (!start) hipStreamWaitEvent(s.stream, s.ev_filled[p], 0);
      launch_batch_sa_sweep(probs, map, nblk, b.d_roff, p, d_betas, d_step, sampler, fg[p], fu[p],
                            s.stream, f32);
      hipEventRecord(s.ev_used[p], s.stream);
      if (!start) hipStreamWaitEvent(s.mt_stream, s.ev_used[1 - p], 0);
      mt_refill_into(s, fg[1 - p], fu[1 - p], s.mt_stream);
      hipEventRecord(s.ev_filled[1 - p], s.mt_stream);
    }
    for (int m = 0; m < nmicro; m++)
      for (int p = 0; p < 2; p++) launch_batch_micro_sweep(probs, map, nblk, p, s.stream, f32);
    launch_batch_advance(d_step, 1, s.stream);
  };
  const auto t0 = std::chrono::steady_clock::now();
  int done = 0;
  if (nsteps >= steps_per_graph) {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    if (hip_check(hipStreamBeginCapture(s.stream, hipStreamCaptureModeThreadLocal), "capture")) rc = GFX_ERR_DEVICE;
    if (rc == GFX_OK) {
      fork();
      for (int k2 = 0; k2 < steps_per_graph; k2++) one_step(k2 == 0);
      join();
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
  if (rc == GFX_OK && done < nsteps) {
    fork();
    for (bool first = true; done < nsteps; first = false, done++) one_step(first);
    join();
  }
  if (f32)
    for (const BProbHost& p : b.probs) launch_g32_to_g64(p.g32, p.g64, p.geo, s.stream);
  if (hip_check(hipStreamSynchronize(s.stream), "anneal") && rc == GFX_OK) rc = GFX_ERR_DEVICE;
  b.run_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  hipFree(d_betas);
  hipFree(d_step);
  return rc;
}

// gfx_batch_anneal in FP64 (GFX_PREC_FP64, default: the reference's
// arithmetic) or on FP32 links and g with FP32 arithmetic (GFX_PREC_MIXED)
int gfx_batch_set_anneal_precision(int mode) {
  if (mode != GFX_PREC_FP64 && mode != GFX_PREC_MIXED) return GFX_ERR_STATE;
  g_b.anneal_fp32 = mode == GFX_PREC_MIXED;
  return GFX_OK;
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
  if (iter) *iter = g_b.probs[q].stop_iter;
  return GFX_OK;
}

int gfx_batch_get_trafo(int q, gfx_su3* trafo) {
  if (q < 0 || q >= (int)g_b.probs.size()) return GFX_ERR_STATE;
  return download_aos_geo(g_st, g_b.probs[q].geo, g_b.probs[q].g64, (double*)trafo, 1);
}

// the g lattice q starts the next gfx_batch_run or gfx_batch_anneal from
int gfx_batch_set_trafo(int q, gfx_su3* trafo) {
  if (q < 0 || q >= (int)g_b.probs.size()) return GFX_ERR_STATE;
  return upload_aos_geo(g_st, g_b.probs[q].geo, (const double*)trafo, g_b.probs[q].g64, 1);
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

int gfx_batch_set_persistent(int on) {
  if (on && !persistent_supported()) return GFX_ERR_DEVICE;
  g_b.persistent = on != 0;
  return GFX_OK;
}

int gfx_batch_persistent_launches(void) { return g_b.persistent_launches; }

// the sweep variant picked for the batch in each precision (index 0 FP64, 1 FP32)
int gfx_batch_variant(int precision, int* variant, double* ms_per_variant) {
  const int pi = precision == GFX_PREC_MIXED ? 1 : 0;
  if (variant) *variant = g_b.variant[pi];
  if (ms_per_variant)
    for (int k = 0; k < SWEEP_VARIANTS; k++) ms_per_variant[k] = g_b.tune_ms[pi][k];
  return g_b.tuned[pi] ? GFX_OK : GFX_ERR_STATE;
}

}  // extern "C"
