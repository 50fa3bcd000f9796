// gfx_api.hip -- the C ABI (include/cudagfx.h): device state of the single
// gauge-fixing instance, host<->device transfers and the overrelaxation
// driver loop (overrelax.cu:677-823) on one HIP stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/cudagfx.h"
#include "gfx_internal.h"
#include "gfx_lattice.h"
#include "gfx_state.h"

using namespace gfx;

extern "C" {
int T = 0, L = 0, LX = 0, LY = 0, LZ = 0, VOLUME = 0;
double FUNC = 0.0, DADA = 0.0, maxDADA = 0.0, PLAQ = 0.0, BETA = 0.0;
}

namespace gfx {
State g_st;

const char* err_text(int code) {
  switch (code) {
    case GFX_OK: return "ok";
    case GFX_ERR_GEOMETRY: return "lattice geometry not supported (all extents must be even)";
    case GFX_ERR_ALLOC: return "device memory allocation failed";
    case GFX_ERR_NOCONV: return "gauge condition not reached";
    case GFX_ERR_WRITE: return "write failed";
    case GFX_ERR_STATE: return "library not initialised for this call";
    case GFX_ERR_DEVICE: return "HIP runtime error";
    case GFX_ERR_MTDAT: return "MersenneTwister.dat missing or unreadable";
    case GFX_ERR_SIZE: return "input larger than the static scratch of the entry point";
    default: return "unknown error";
  }
}

int hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    fprintf(stderr, "cudagfx: %s: %s\n", what, hipGetErrorString(e));
    return GFX_ERR_DEVICE;
  }
  return GFX_OK;
}

Geom make_geom(int l0, int l1, int l2, int l3) {
  Geom g;
  g.L[0] = l0;
  g.L[1] = l1;
  g.L[2] = l2;
  g.L[3] = l3;
  g.h0 = l0 / 2;
  g.nh = (int)((long)l0 * l1 * l2 * l3 / 2);
  g.st[0] = 1;
  g.st[1] = g.h0;
  g.st[2] = g.h0 * l1;
  g.st[3] = g.h0 * l1 * l2;
  // planes 64-element aligned; a power-of-two plane distance puts the six
  // planes of a site on the same HBM channel, so break it
  int s = (g.nh + 63) & ~63;
  if ((s & (s - 1)) == 0) s += 64;
  g.stride = s;
  g.dh0 = make_fastdiv(g.h0);
  g.dL1 = make_fastdiv(l1);
  g.dL2 = make_fastdiv(l2);
  return g;
}

// even extents, at most 2^29 sites (2^28 per parity: the kernels address a
// plane's elements with 32-bit byte offsets; the FP64 links of 2^29 sites
// alone take 206 GB)
bool geometry_ok(int l0, int l1, int l2, int l3) {
  const int d[4] = {l0, l1, l2, l3};
  long v = 1;
  for (int k = 0; k < 4; k++) {
    if (d[k] < 2 || (d[k] & 1)) return false;
    v *= d[k];
  }
  return v <= (1L << 29);
}

int ensure_common(int need_links32) {
  State& s = g_st;
  if (!s.stream) {
    if (hip_check(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking), "stream")) return GFX_ERR_DEVICE;
    hipEventCreate(&s.ev0);
    hipEventCreate(&s.ev1);
  }
  (void)need_links32;
  return GFX_OK;
}

int dev_alloc(void** p, size_t bytes, const char* what) {
  if (hipMalloc(p, bytes) != hipSuccess) {
    fprintf(stderr, "cudagfx: allocation of %s (%zu bytes) failed\n", what, bytes);
    *p = nullptr;
    return GFX_ERR_ALLOC;
  }
  return GFX_OK;
}

size_t field_bytes64(const Geom& g) { return (size_t)6 * g.stride * sizeof(double2); }
size_t field_bytes32(const Geom& g) { return (size_t)3 * g.stride * sizeof(float4); }

bool block_order(const Geom& geo, int block, int tile, std::vector<int>& order) {
  const int nblk = (geo.nh + block - 1) / block;
  order.resize(nblk);
  for (int b = 0; b < nblk; b++) order[b] = b;
  if (tile <= 0) return true;
  const int nrow = geo.h0 * geo.L[1];  // sites of one parity in an (x2, x3) plane
  const int l2 = geo.L[2], l3 = geo.L[3];
  int k = 0;
  if (tile == 3) {
    // chunks in x1 as well: pieces of C1 blocks of a plane x C2 planes in
    // x2, x3 outermost inside a chunk (layers of about 4096 sites)
    if (nrow % block != 0) return false;
    const int bpp = nrow / block;  // blocks per plane
    const int C1 = std::min(bpp, std::max(1, 512 / block));
    const int C2 = std::min(l2, 8);
    if (C1 == bpp && C2 == l2) return false;  // one chunk: lexicographic
    for (int c1 = 0; c1 < bpp; c1 += C1)
      for (int c2 = 0; c2 < l2; c2 += C2)
        for (int x3 = 0; x3 < l3; x3++)
          for (int x2 = c2; x2 < std::min(c2 + C2, l2); x2++)
            for (int b = c1; b < std::min(c1 + C1, bpp); b++) order[k++] = b + bpp * (x2 + l2 * x3);
    return k == nblk;
  }
  int a = nrow, c = block;
  while (c) {
    const int t = a % c;
    a = c;
    c = t;
  }
  const int G = block / a;  // rows per superrow: G * nrow is a multiple of block
  if (l2 % G != 0) return false;
  const int ns2 = l2 / G;
  const int bps = G * nrow / block;  // blocks per superrow
  const int C = 2 << tile;           // tile 1: 4 superrows, tile 2: 8
  if (C >= ns2) return false;  // one chunk: the lexicographic order
  for (int c0 = 0; c0 < ns2; c0 += C) {
    const int cw = std::min(C, ns2 - c0);
    for (int x3 = 0; x3 < l3; x3++)
      for (int x2s = c0; x2s < c0 + cw; x2s++)
        for (int sub = 0; sub < bps; sub++) order[k++] = (x2s + ns2 * x3) * bps + sub;
  }
  return k == nblk;
}

int alloc_reduction(State& s) {
  const int nb = functional_blocks(s.geo.nh) + 8;
  // partials: max over Landau (1 segment) and Coulomb (T segments)
  const int nbc = functional_blocks(s.geo.nh / s.geo.L[3]) * s.geo.L[3] + 8;
  s.npartials = (nb > nbc ? nb : nbc);
  if (dev_alloc((void**)&s.partials, sizeof(double) * 3 * s.npartials, "partials")) return GFX_ERR_ALLOC;
  if (dev_alloc((void**)&s.results, sizeof(double) * 4 * (s.geo.L[3] + 1), "results")) return GFX_ERR_ALLOC;
  if (hipHostMalloc((void**)&s.h_results, sizeof(double) * 4 * (s.geo.L[3] + 1), 0) != hipSuccess) return GFX_ERR_ALLOC;
  // per time slice: sweeping in FP64 [0, T), sweeping in FP32 [T, 2T)
  if (dev_alloc((void**)&s.active, 2 * (size_t)s.geo.L[3], "active")) return GFX_ERR_ALLOC;
  s.ncheck = (s.geo.nh + 63) / 64;  // blocks of the smallest sweep workgroup
  // check partials: odd + even sweeps, for each precision (Coulomb runs both)
  if (dev_alloc((void**)&s.cpart, sizeof(double) * 12 * s.ncheck, "check partials")) return GFX_ERR_ALLOC;
  for (int pr = 0; pr < 2; pr++) {
    if (dev_alloc((void**)&s.smap[pr], sizeof(int) * s.ncheck, "slice map")) return GFX_ERR_ALLOC;
    s.smap_n[pr] = 0;
  }
  for (int bi = 0; bi < 2; bi++)
    for (int t = 0; t < TILINGS; t++) {
      std::vector<int> order, map;
      s.tiling_ok[bi][t] = block_order(s.geo, bi == 0 ? 256 : 64, t, order);
      deal_to_xcds(order, map);
      if (dev_alloc((void**)&s.bmap[bi][t], sizeof(int) * map.size(), "workgroup map")) return GFX_ERR_ALLOC;
      if (hip_check(hipMemcpy(s.bmap[bi][t], map.data(), sizeof(int) * map.size(), hipMemcpyHostToDevice), "workgroup map"))
        return GFX_ERR_DEVICE;
    }
  return GFX_OK;
}

// staging chunk for AoS <-> SoA conversions (sites per chunk)
static const int CHUNK = 32768;

int ensure_staging(State& s) {
  if (s.staging) return GFX_OK;
  return dev_alloc((void**)&s.staging, (size_t)CHUNK * 4 * 18 * sizeof(double), "staging");
}

// host AoS su3 (nmat per site) -> device FP64 fields of lattice geo
int upload_aos_geo(State& s, const Geom& geo, const double* host, double2* dst, int nmat) {
  if (ensure_staging(s)) return GFX_ERR_ALLOC;
  const int V = 2 * geo.nh;
  for (int s0 = 0; s0 < V; s0 += CHUNK) {
    const int ns = (V - s0 < CHUNK) ? V - s0 : CHUNK;
    const size_t bytes = (size_t)ns * nmat * 18 * sizeof(double);
    if (hip_check(hipMemcpyAsync(s.staging, host + (size_t)s0 * nmat * 18, bytes, hipMemcpyHostToDevice, s.stream), "H2D")) return GFX_ERR_DEVICE;
    launch_pack_f64(s.staging, dst, geo, s0, ns, nmat, s.stream);
  }
  return hip_check(hipStreamSynchronize(s.stream), "upload");
}

int download_aos_geo(State& s, const Geom& geo, const double2* src, double* host, int nmat) {
  if (ensure_staging(s)) return GFX_ERR_ALLOC;
  const int V = 2 * geo.nh;
  for (int s0 = 0; s0 < V; s0 += CHUNK) {
    const int ns = (V - s0 < CHUNK) ? V - s0 : CHUNK;
    const size_t bytes = (size_t)ns * nmat * 18 * sizeof(double);
    launch_unpack_f64(src, s.staging, geo, s0, ns, nmat, s.stream);
    if (hip_check(hipMemcpyAsync(host + (size_t)s0 * nmat * 18, s.staging, bytes, hipMemcpyDeviceToHost, s.stream), "D2H")) return GFX_ERR_DEVICE;
    if (hip_check(hipStreamSynchronize(s.stream), "download")) return GFX_ERR_DEVICE;
  }
  return GFX_OK;
}

int upload_aos(State& s, const double* host, double2* dst, int nmat) {
  return upload_aos_geo(s, s.geo, host, dst, nmat);
}
int download_aos(State& s, const double2* src, double* host, int nmat) {
  return download_aos_geo(s, s.geo, src, host, nmat);
}

int ensure_f32_links(State& s) {
  if (s.U32 && s.U32_valid) return GFX_OK;
  if (!s.U32) {
    if (dev_alloc((void**)&s.U32, 8 * field_bytes32(s.geo), "FP32 links")) return GFX_ERR_ALLOC;
    if (dev_alloc((void**)&s.g32, 4 * field_bytes32(s.geo), "FP32 trafo")) return GFX_ERR_ALLOC;
  }
  launch_f64_to_f32(s.U64, s.U32, s.geo, 8, s.stream);
  s.U32_valid = true;
  s.packed_valid = false;
  s.fx_valid = false;
  return GFX_OK;
}

// the packed links of the mixed runs (92 bytes per link in one allocation:
// a, b, c planes of the 24-bit high parts, the low parts, then the two
// planes of the 21-bit links); also makes the FP32 fields, whose validity
// they share
int ensure_packed_links(State& s) {
  int rc = ensure_f32_links(s);
  if (rc) return rc;
  if (s.packed_valid) return GFX_OK;
  if (!s.pk.a) {
    const size_t n = (size_t)8 * s.geo.stride;
    char* base = nullptr;
    if (dev_alloc((void**)&base, n * (16 + 16 + 4 + 24 + 32), "packed links")) return GFX_ERR_ALLOC;
    s.pk.a = reinterpret_cast<uint4*>(base);
    s.pk.b = reinterpret_cast<uint4*>(base + n * 16);
    s.pk.c = reinterpret_cast<uint32_t*>(base + n * 32);
    s.pk.lo = reinterpret_cast<uint2*>(base + n * 36);
    s.pk.a21 = reinterpret_cast<uint4*>(base + n * 60);
    s.pk.b21 = reinterpret_cast<uint4*>(base + n * 76);
  }
  launch_pack_links(s.U64, s.pk, s.geo, 8, s.stream);
  s.packed_valid = true;
  return GFX_OK;
}

size_t field_bytes_fx(const Geom& g) { return (size_t)3 * g.stride * sizeof(int4); }

// the 32-bit fixed-point links and g of the mixed Landau run's polish;
// the links share the validity of the FP32 fields
int ensure_fx_links(State& s) {
  int rc = ensure_f32_links(s);
  if (rc) return rc;
  if (s.fx_valid) return GFX_OK;
  if (!s.Ufx) {
    if (dev_alloc((void**)&s.Ufx, 8 * field_bytes_fx(s.geo), "fixed-point links")) return GFX_ERR_ALLOC;
    if (dev_alloc((void**)&s.gfx, 4 * field_bytes_fx(s.geo), "fixed-point trafo")) return GFX_ERR_ALLOC;
  }
  launch_f64_to_fx(s.U64, s.Ufx, s.geo, 8, s.stream);
  s.fx_valid = true;
  return GFX_OK;
}

void free_all(State& s) {
  if (s.stream) hipStreamSynchronize(s.stream);
  hipFree(s.U64);
  hipFree(s.g64);
  hipFree(s.U32);
  hipFree(s.g32);
  hipFree(s.pk.a);
  s.pk = PackedArgs{};
  s.packed_valid = false;
  hipFree(s.Ufx);
  hipFree(s.gfx);
  s.Ufx = nullptr;
  s.gfx = nullptr;
  s.fx_valid = false;
  hipFree(s.partials);
  hipFree(s.cpart);
  for (int bi = 0; bi < 2; bi++)
    for (int t = 0; t < TILINGS; t++) {
      hipFree(s.bmap[bi][t]);
      s.bmap[bi][t] = nullptr;
    }
  for (int pr = 0; pr < 2; pr++) {
    hipFree(s.smap[pr]);
    s.smap[pr] = nullptr;
    s.smap_n[pr] = 0;
  }
  hipFree(s.results);
  hipFree(s.active);
  hipFree(s.staging);
  if (s.h_results) hipHostFree(s.h_results);
  s.U64 = nullptr;
  s.g64 = nullptr;
  s.U32 = nullptr;
  s.g32 = nullptr;
  s.partials = nullptr;
  s.cpart = nullptr;
  s.results = nullptr;
  s.active = nullptr;
  s.staging = nullptr;
  s.h_results = nullptr;
  s.U32_valid = false;
  s.init = 0;
}

int set_lattice_from_globals(State& s) {
  if (!geometry_ok(LX, LY, LZ, T)) {
    printf("Error: lattice %d x %d x %d x %d not supported (extents must be even, at most 2^29 sites). Aborting...\n",
           LX, LY, LZ, T);
    return GFX_ERR_GEOMETRY;
  }
  VOLUME = LX * LY * LZ * T;
  s.geo = make_geom(LX, LY, LZ, T);
  return GFX_OK;
}

// one functional evaluation into s.h_results[0..2] for nseg segments
int functional_now(State& s, bool fp32, int ndir, int nseg, double eps, int even, int odd) {
  const size_t S = s.geo.stride;
  if (fp32)
    launch_functional_f32(s.U32, s.g32 + (size_t)even * 3 * S, s.g32 + (size_t)odd * 3 * S, s.geo,
                          ndir, nseg, s.partials, eps, s.results, s.stream);
  else
    launch_functional_f64(s.U64, s.g64 + (size_t)even * 6 * S, s.g64 + (size_t)odd * 6 * S, s.geo,
                          ndir, nseg, s.partials, eps, s.results, s.stream);
  if (hip_check(hipMemcpyAsync(s.h_results, s.results, sizeof(double) * 4 * nseg, hipMemcpyDeviceToHost, s.stream), "D2H results")) return GFX_ERR_DEVICE;
  return hip_check(hipStreamSynchronize(s.stream), "functional");
}

void reset_tuning(State& s) {
  s.hb_tiling = -1;
  s.sa_tiling = -1;
  for (int a = 0; a < NFORM; a++)
    for (int b = 0; b < 2; b++) {
      s.tuned[a][b] = false;
      s.variant[a][b] = 0;
      for (int k = 0; k < SWEEP_VARIANTS; k++) s.tune_ms[a][b][k] = 0.f;
    }
}

// block partials of the fused checks of one precision: odd sweep at +0,
// even sweep at +3 ncheck
double* check_partials(State& s, bool fp32) { return s.cpart + (fp32 ? 6 * (size_t)s.ncheck : 0); }

const int* sweep_map(State& s, int v) {
  return s.bmap[block_index(sweep_variant_block(v))][sweep_variant_tiling(v)];
}

double2* field64(State& s, int k) { return s.g64 + (size_t)k * 6 * s.geo.stride; }
float4* field32(State& s, int k) { return s.g32 + (size_t)k * 3 * s.geo.stride; }
int4* fieldfx(State& s, int k) { return s.gfx + (size_t)k * 3 * s.geo.stride; }

// the form of the next sweep: FP32 or FP64, on the packed links while a
// mixed run reads them
int sweep_form(const State& s, bool fp32, int ndir) {
  (void)ndir;
  if (s.packed_on) return fp32 ? FORM_P32 : s.fx_on ? FORM_X64 : FORM_P64;
  return fp32 ? FORM_F32 : FORM_F64;
}

// one half sweep with the tuned variant for (form, ndir) on the even field
// in slot `even` (0 or 2) and the odd field in slot `odd` (1 or 3): in
// place, or with chk = 2 (as in launch_sweep_f64) the new even field into
// slot 2 - even, or with `spare` the new odd field into slot 4 - odd; the
// block partials of a check go to cpart (odd sweep first half, even sweep
// second half)
static void close_sample(State& s) {
  hipEventRecord(s.sample_ev[2 * s.sample_n + 1], s.stream);
  s.sample_open = false;
  s.sample_n++;
}

void do_sweep(State& s, bool fp32, int parity, int ndir, const unsigned char* active, int even,
              int chk, int odd = 1, bool spare = false) {
  const int form = sweep_form(s, fp32, ndir);
  const int v = s.variant[form][ndir == 4 ? 1 : 0];
  double* part = check_partials(s, fp32) + (chk == 2 ? 3 * (size_t)s.ncheck : 0);
  const int* map = sweep_map(s, v);
  int nb = 0;
  if (ndir == 3 && s.smap_n[fp32 ? 1 : 0] > 0) {
    map = s.smap[fp32 ? 1 : 0];
    nb = s.smap_n[fp32 ? 1 : 0];
  }
  // output slot of the updated parity (-1: in place)
  const int out = (chk == 2) ? 2 - even : spare ? 4 - odd : -1;
  if (s.sample_on) {
    // a check sweep, a sweep run ahead of a check's result or another form
    // ends the open segment; a plain sweep opens one if none is open
    if (s.sample_open && (chk != 0 || spare || s.sample_form[s.sample_n] != form)) close_sample(s);
    if (!s.sample_open && chk == 0 && !spare && s.sample_n < State::SAMPLE_MAX) {
      hipEventRecord(s.sample_ev[2 * s.sample_n], s.stream);
      s.sample_form[s.sample_n] = form;
      s.sample_cnt[s.sample_n] = 0;
      s.sample_open = true;
    }
  }
  if (form == FORM_F32 || form == FORM_P32) {
    float4* alt = out < 0 ? nullptr : field32(s, out);
    if (form == FORM_F32)
      launch_sweep_f32(s.U32, field32(s, even), field32(s, odd), s.geo, parity, ndir, s.omega,
                       active, v, map, s.stream, chk, alt, part, nb);
    else
      launch_sweep_packed_f32(s.pk, field32(s, even), field32(s, odd), s.geo, parity, ndir,
                              s.omega, active, v, map, s.stream, chk, alt, part, nb);
  } else {
    double2* alt = out < 0 ? nullptr : field64(s, out);
    if (form == FORM_X64)
      launch_sweep_fx(s.Ufx, fieldfx(s, even), fieldfx(s, odd), s.geo, parity, ndir, s.omega,
                      active, v, map, s.stream, chk, out < 0 ? nullptr : fieldfx(s, out), part, nb);
    else if (form == FORM_P64)
      launch_sweep_packed_f64(s.pk, field64(s, even), field64(s, odd), s.geo, parity, ndir,
                              s.omega, active, v, map, s.stream, chk, alt, part, nb);
    else
      launch_sweep_f64(s.U64, field64(s, even), field64(s, odd), s.geo, parity, ndir, s.omega,
                       active, v, map, s.stream, chk, alt, part, nb);
  }
  if (s.sample_open) s.sample_cnt[s.sample_n]++;
}
void do_sweep(State& s, bool fp32, int parity, int ndir, const unsigned char* active) {
  do_sweep(s, fp32, parity, ndir, active, 0, 0);
}

static void stop_clock_ms(State& s, std::chrono::steady_clock::time_point t0) {
  hipStreamSynchronize(s.stream);
  s.run_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// Coulomb: the map of precision fp32's variant restricted to the blocks
// that hold a site of a slice still being fixed (act), dealt to the XCDs
static void build_slice_map(State& s, bool fp32, const std::vector<unsigned char>& act, int nsh) {
  const int pr = fp32 ? 1 : 0;
  const int v = s.variant[sweep_form(s, fp32, 3)][0];
  const int blk = sweep_variant_block(v);
  std::vector<int> order, list;
  block_order(s.geo, blk, sweep_variant_tiling(v), order);
  list.reserve(order.size());
  for (int b : order) {
    const int t0 = (b * blk) / nsh;
    const int t1 = (std::min((b + 1) * blk, s.geo.nh) - 1) / nsh;
    bool any = false;
    for (int t = t0; t <= t1 && !any; t++) any = act[t] != 0;
    if (any) list.push_back(b);
  }
  deal_to_xcds(list, s.smap_h[pr]);
  s.smap_n[pr] = (int)s.smap_h[pr].size();
  if (s.smap_n[pr] > 0)
    hipMemcpyAsync(s.smap[pr], s.smap_h[pr].data(), sizeof(int) * s.smap_n[pr],
                   hipMemcpyHostToDevice, s.stream);
}

// bring the even field back to slot 0 (the odd field to slot 1)
static void canonical_slot(State& s, bool fp32, int slot, int home, bool fx = false) {
  if (slot == home) return;
  if (fx)
    hipMemcpyAsync(fieldfx(s, home), fieldfx(s, slot), field_bytes_fx(s.geo), hipMemcpyDeviceToDevice, s.stream);
  else if (fp32)
    hipMemcpyAsync(field32(s, home), field32(s, slot), field_bytes32(s.geo), hipMemcpyDeviceToDevice, s.stream);
  else
    hipMemcpyAsync(field64(s, home), field64(s, slot), field_bytes64(s.geo), hipMemcpyDeviceToDevice, s.stream);
}
static void canonical_even(State& s, bool fp32, int even) { canonical_slot(s, fp32, even, 0); }

// Times every sweep variant on a scratch copy of g (the state is not
// touched) and keeps the fastest for this lattice, form and ndir.  The
// packed FP64 form has rolled variants only.
int tune_sweep(State& s, int form, int ndir) {
  const int pi = form, di = ndir == 4 ? 1 : 0;
  if (!s.autotune || s.tuned[pi][di]) return GFX_OK;
  const bool fp32 = form == FORM_F32 || form == FORM_P32;
  const size_t bytes = 2 * (form == FORM_X64 ? field_bytes_fx(s.geo)
                            : fp32           ? field_bytes32(s.geo)
                                             : field_bytes64(s.geo));
  void* scratch = nullptr;
  if (hipMalloc(&scratch, bytes) != hipSuccess) {
    s.tuned[pi][di] = true;  // keep the default variant
    return GFX_OK;
  }
  hipMemcpyAsync(scratch, form == FORM_X64 ? (void*)s.gfx : fp32 ? (void*)s.g32 : (void*)s.g64, bytes,
                 hipMemcpyDeviceToDevice, s.stream);
  float4* sc32 = static_cast<float4*>(scratch);
  double2* sc64 = static_cast<double2*>(scratch);
  int4* scfx = static_cast<int4*>(scratch);
  auto sweep = [&](int v, const int* map, int p) {
    if (form == FORM_X64)
      launch_sweep_fx(s.Ufx, scfx, scfx + 3 * (size_t)s.geo.stride, s.geo, p, ndir, s.omega, nullptr, v,
                      map, s.stream);
    else if (form == FORM_F32)
      launch_sweep_f32(s.U32, sc32, sc32 + 3 * (size_t)s.geo.stride, s.geo, p, ndir, s.omega, nullptr, v, map, s.stream);
    else if (form == FORM_P32)
      launch_sweep_packed_f32(s.pk, sc32, sc32 + 3 * (size_t)s.geo.stride, s.geo, p, ndir, s.omega, nullptr, v, map, s.stream);
    else if (form == FORM_P64)
      launch_sweep_packed_f64(s.pk, sc64, sc64 + 6 * (size_t)s.geo.stride, s.geo, p, ndir, s.omega, nullptr, v, map, s.stream);
    else
      launch_sweep_f64(s.U64, sc64, sc64 + 6 * (size_t)s.geo.stride, s.geo, p, ndir, s.omega, nullptr, v, map, s.stream);
  };
  float best = 1e30f;
  int bestv = s.variant[pi][di];
  for (int v = 0; v < SWEEP_VARIANTS; v++) {
    if (!s.tiling_ok[block_index(sweep_variant_block(v))][sweep_variant_tiling(v)]) continue;
    if (form == FORM_P64 && !sweep_variant_rolled(v)) continue;
    const int* map = sweep_map(s, v);
    for (int rep = 0; rep < 2; rep++)
      for (int p = 0; p < 2; p++) sweep(v, map, p);
    // 6 timed sweeps: the fastest variants of a form lie within 2-3% of
    // each other, about the spread of a 3-sweep timing
    hipEventRecord(s.ev0, s.stream);
    for (int rep = 0; rep < 6; rep++)
      for (int p = 0; p < 2; p++) sweep(v, map, p);
    hipEventRecord(s.ev1, s.stream);
    if (hip_check(hipEventSynchronize(s.ev1), "tune")) break;
    float t = 0.f;
    hipEventElapsedTime(&t, s.ev0, s.ev1);
    t /= 12.0f;
    s.tune_ms[pi][di][v] = t;
    if (t < best) {
      best = t;
      bestv = v;
    }
  }
  hipStreamSynchronize(s.stream);
  hipFree(scratch);
  s.variant[pi][di] = bestv;
  s.tuned[pi][di] = true;
  return GFX_OK;
}

}  // namespace gfx

// ======================================================================
extern "C" {

const char* gfx_error_string(int code) { return err_text(code); }

int init_gaugefixing(gfx_su3* gf, gfx_su3* trafo) {
  State& s = g_st;
  if (s.init) free_all(s);
  int rc = set_lattice_from_globals(s);
  if (rc) return rc;
  if ((rc = ensure_common(0))) return rc;
  if (dev_alloc((void**)&s.U64, 8 * field_bytes64(s.geo), "gauge field")) return GFX_ERR_ALLOC;
  printf("Allocated gauge field on device\n");
  if (dev_alloc((void**)&s.g64, 4 * field_bytes64(s.geo), "trafo field")) return GFX_ERR_ALLOC;
  printf("Allocated trafo field 1 on device\n");
  if ((rc = alloc_reduction(s))) return rc;
  if ((rc = upload_aos(s, (const double*)gf, s.U64, 4))) return rc;
  if ((rc = upload_aos(s, (const double*)trafo, s.g64, 1))) return rc;
  s.U32_valid = false;
  reset_tuning(s);
  s.init = 1;
  return GFX_OK;
}

void finalize_gaugefixing(void) { free_all(g_st); }

int init_thermalization(gfx_su3* gf) {
  State& s = g_st;
  if (s.init) free_all(s);
  int rc = set_lattice_from_globals(s);
  if (rc) return rc;
  if ((rc = ensure_common(0))) return rc;
  if (dev_alloc((void**)&s.U64, 8 * field_bytes64(s.geo), "gauge field")) return GFX_ERR_ALLOC;
  printf("Allocated gauge field on device\n");
  if (dev_alloc((void**)&s.g64, 4 * field_bytes64(s.geo), "trafo field")) return GFX_ERR_ALLOC;
  if ((rc = alloc_reduction(s))) return rc;
  if ((rc = upload_aos(s, (const double*)gf, s.U64, 4))) return rc;
  s.U32_valid = false;
  reset_tuning(s);
  s.init = 2;
  return GFX_OK;
}

void finalize_thermalization(void) { free_all(g_st); }

int gfx_set_omega(double omega) {
  g_st.omega = omega;
  return GFX_OK;
}
int gfx_set_precision(int mode, double switch_theta) {
  if (mode != GFX_PREC_FP64 && mode != GFX_PREC_MIXED) return GFX_ERR_STATE;
  g_st.precision = mode;
  if (switch_theta > 0) g_st.switch_theta = switch_theta;
  return GFX_OK;
}
int gfx_set_packed_links(int on) {
  if (on != 0 && on != 1) return GFX_ERR_STATE;
  g_st.packed_links = on;
  return GFX_OK;
}
int gfx_set_fp32_stall(int iters, double gate) {
  if (iters < 0 || !(gate > 0)) return GFX_ERR_STATE;
  g_st.stall_iters = iters;
  g_st.stall_gate = gate;
  return GFX_OK;
}
int gfx_set_verbose(int v) {
  g_st.verbose = v;
  return GFX_OK;
}

double calc_functional(void) {
  State& s = g_st;
  if (!s.init) return -1.0;
  if (functional_now(s, false, 4, 1, 0.0)) return -1.0;
  FUNC = s.h_results[0];
  DADA = s.h_results[1];
  maxDADA = s.h_results[2];
  return maxDADA;
}

int gfx_functional_slices(double* F, double* D, double* M) {
  State& s = g_st;
  if (!s.init) return GFX_ERR_STATE;
  const int nseg = s.geo.L[3];
  if (functional_now(s, false, 3, nseg, 0.0)) return GFX_ERR_DEVICE;
  for (int t = 0; t < nseg; t++) {
    if (F) F[t] = s.h_results[4 * t];
    if (D) D[t] = s.h_results[4 * t + 1];
    if (M) M[t] = s.h_results[4 * t + 2];
  }
  return GFX_OK;
}

double calc_plaquette(int outputyn) {
  State& s = g_st;
  if (!s.init) return -1.0;
  launch_plaquette_f64(s.U64, s.geo, s.partials, s.results, s.stream);
  hipMemcpyAsync(s.h_results, s.results, sizeof(double), hipMemcpyDeviceToHost, s.stream);
  if (hip_check(hipStreamSynchronize(s.stream), "plaquette")) return -1.0;
  const double p = s.h_results[0];
  if (outputyn == 1) printf("PLAQ = %.16f\n", p);
  return p;
}

int gfx_halfsweep(int parity, int ndir) {
  State& s = g_st;
  if (!s.init) return GFX_ERR_STATE;
  if (s.precision == GFX_PREC_MIXED) {
    int rc = ensure_f32_links(s);
    if (rc) return rc;
    launch_f64_to_f32(s.g64, s.g32, s.geo, 2, s.stream);
    do_sweep(s, true, parity, ndir, nullptr);
    launch_g32_to_g64(s.g32, s.g64, s.geo, s.stream);
  } else {
    do_sweep(s, false, parity, ndir, nullptr);
  }
  return hip_check(hipStreamSynchronize(s.stream), "halfsweep");
}

int overrelax_gauge(int maxit, double eps, int checkinterval) {
  State& s = g_st;
  if (!s.init) return -1;
  if (checkinterval < 1) checkinterval = 1;
  bool fp32 = (s.precision == GFX_PREC_MIXED);
  const bool packed = fp32 && s.packed_links;
  // the polish of a run on packed links holds links and g in 32-bit fixed
  // point, for eps well above what that resolves (below it: high + low parts
  // and FP64 g)
  const bool fx_polish = packed && eps > FX_EPS_MIN;
  // every exit leaves the packed forms off (gfx_halfsweep and the other
  // entry points sweep on the FP32 and FP64 links)
  struct PackedOff {
    State& s;
    ~PackedOff() {
      s.packed_on = false;
      s.fx_on = false;
    }
  } packed_off{s};
  s.packed_on = false;
  s.fx_on = false;
  if (fp32) {
    if (ensure_f32_links(s)) return -1;
    launch_f64_to_f32(s.g64, s.g32, s.geo, 2, s.stream);
  }
  if (packed) {
    if (ensure_packed_links(s)) return -1;
    if (fx_polish && ensure_fx_links(s)) return -1;
    tune_sweep(s, FORM_P32, 4);
    if (fx_polish) {
      launch_f64_to_fx(s.g64, s.gfx, s.geo, 2, s.stream);
      tune_sweep(s, FORM_X64, 4);
    } else {
      tune_sweep(s, FORM_P64, 4);
    }
  } else if (fp32) {
    tune_sweep(s, FORM_F32, 4);
  }
  tune_sweep(s, FORM_F64, 4);
  s.packed_on = packed;
  s.run_iters32 = 0;
  s.run_iters64 = 0;
  if (s.sample_open) close_sample(s);
  s.sample_n = 0;
  const double V = (double)VOLUME;
  const int stall_checks = (s.stall_iters + checkinterval - 1) / checkinterval;
  double maxdada = 0.0, best = 1e300;
  int stall = 0;
  int even = 0;            // slot of the current even field (0 or 2)
  int odd = 1;             // slot of the current odd field (1 or 3)
  bool even_done = false;  // the even half of iteration i ran with the last check
  bool odd_done = false;   // and its odd half too (into the spare odd slot)
  bool converged = false;
  int i;
  hipStreamSynchronize(s.stream);
  const auto t0 = std::chrono::steady_clock::now();
  // the polish's g is in the fixed-point fields (gfx) while fx
  bool fx = false;
  // state i (the checked state) back into slots 0 and 1 of the fields the
  // run sweeps now
  auto canonical = [&]() {
    canonical_slot(s, fp32, even, 0, fx);
    canonical_slot(s, fp32, odd, 1, fx);
    even = 0;
    odd = 1;
  };
  // state i into the FP64 fields g64 (reunitarised in FP64 when it comes
  // from FP32 or fixed-point g), the run's sweeps then plain FP64
  auto to_fp64 = [&]() {
    canonical();
    if (fp32)
      launch_g32_to_g64(s.g32, s.g64, s.geo, s.stream);
    else if (fx)
      launch_fx_to_g64(s.gfx, s.g64, s.geo, s.stream);
    fp32 = fx = s.fx_on = false;
  };
  // the FP64 state in g64 into the fields of the polish
  auto to_polish = [&]() {
    if (fx_polish) {
      launch_f64_to_fx(s.g64, s.gfx, s.geo, 2, s.stream);
      fx = s.fx_on = true;
    }
  };
  for (i = 0; i < maxit; i++) {
    const bool chk = (i % checkinterval) == 0;
    if (!even_done) do_sweep(s, fp32, 0, 4, nullptr, even, 0, odd);
    if (!odd_done) do_sweep(s, fp32, 1, 4, nullptr, even, chk ? 1 : 0, odd);
    even_done = odd_done = false;
    if (fp32) s.run_iters32++; else s.run_iters64++;
    if (!chk) continue;
    // the check of state i rides on the even sweep of iteration i+1
    do_sweep(s, fp32, 0, 4, nullptr, even, 2, odd);
    const int v = s.variant[sweep_form(s, fp32, 4)][1];
    const double* pc = check_partials(s, fp32);
    launch_check_finalize(pc, pc + 3 * (size_t)s.ncheck, sweep_blocks(s.geo.nh, v), 4.0 * V, V,
                          eps, s.results, s.stream);
    if (hip_check(hipMemcpyAsync(s.h_results, s.results, sizeof(double) * 4, hipMemcpyDeviceToHost, s.stream), "D2H results") ||
        hip_check(hipEventRecord(s.ev1, s.stream), "check event"))
      return -1;
    // while the result travels to the host, the GPU runs the odd sweep of
    // iteration i+1 into the spare odd slot: kept if the run goes on as it
    // is, dropped (state i is untouched) if it stops or changes form here
    const bool ahead = i + 1 < maxit;
    if (ahead) do_sweep(s, fp32, 1, 4, nullptr, 2 - even, ((i + 1) % checkinterval) == 0 ? 1 : 0, odd, true);
    if (hip_check(hipEventSynchronize(s.ev1), "check")) return -1;
    FUNC = s.h_results[0];
    DADA = s.h_results[1];
    maxDADA = maxdada = s.h_results[2];
    if (s.verbose)
      printf("iter %6d:\t FUNC = %.16e \t dAdA = %.16e\t max(dAdA) = %.16e\n", i, FUNC, DADA, maxdada);
    if (maxdada < eps) {
      if (fp32 || s.packed_on) {
        // a stop found by an FP32 sweep or by the polish on the packed
        // links is decided on the FP64 links: state i (converted to FP64 g
        // with an FP64 reunitarisation if it is an FP32 state) with the
        // per-link functional
        const bool from32 = fp32;
        to_fp64();
        if (functional_now(s, false, 4, 1, eps)) return -1;
        FUNC = s.h_results[0];
        DADA = s.h_results[1];
        maxDADA = maxdada = s.h_results[2];
        if (s.verbose)
          printf("iter %6d:\t FUNC = %.16e \t dAdA = %.16e\t max(dAdA) = %.16e (FP64 links)\n", i,
                 FUNC, DADA, maxdada);
        if (!(maxdada < eps)) {
          // not there on the FP64 links: iteration i+1 onwards in FP64, the
          // polish form after an FP32 stop, plain FP64 after a polish stop
          if (from32)
            to_polish();
          else
            s.packed_on = false;
          continue;
        }
      }
      converged = true;
      break;
    }
    if (fp32) {
      bool stalled = false;
      if (maxdada < 0.5 * best) {
        best = maxdada;
        stall = 0;
      } else if (best < s.stall_gate && ++stall >= stall_checks) {
        stalled = true;
      }
      if (maxdada < s.switch_theta || stalled) {
        // leave FP32 at the checked state: g reunitarised in FP64
        if (fx_polish) {
          canonical();
          launch_g32_to_fx(s.g32, s.gfx, s.geo, s.stream);
          fp32 = false;
          fx = s.fx_on = true;
        } else {
          to_fp64();
        }
        continue;
      }
    }
    if (ahead) {
      even = 2 - even;
      odd = 4 - odd;
      even_done = odd_done = true;
    }
  }
  if (s.sample_open) close_sample(s);
  to_fp64();
  s.packed_on = false;
  stop_clock_ms(s, t0);
  if (converged) {
    if (s.verbose) {
      printf("CONVERGENCE!\nFinal values: F = %.16e dAdA = %.16e, maxdAdA = %.16e\n", FUNC, DADA, maxDADA);
      printf("Overrelaxation finished after %f sec\n", s.run_ms / 1000.0);
    }
    return i;
  }
  if (s.verbose) printf("Overrelaxation finished after %f sec\n", s.run_ms / 1000.0);
  if (maxdada > eps) {
    if (s.verbose)
      printf("FAIL! Gauge condition not reached!\nFinal values: F = %.16e dAdA = %.16e, maxdAdA = %.16e\n", FUNC, DADA, maxDADA);
    return -1;
  }
  return i;
}

// Coulomb gauge: every time slice is fixed with the 3 spatial directions and
// stops on its own (per-slice max dAdA < eps, checked every checkinterval
// iterations); a stopped slice is no longer updated, and the sweeps launch
// only the blocks of the slices still being fixed.  In the mixed form every
// slice starts in FP32 and moves to FP64 on its own (per-slice max dAdA <
// switch, or its FP32 stall), so FP32 and FP64 slices are swept side by
// side.  When the slices hold a whole number of sweep blocks the checks ride
// on the sweeps as for Landau (per-slice partials); a slice that stops keeps
// the state of its check: its even sites are copied back over the check
// sweep's new values before the fields swap.
int gfx_coulomb_gauge(int maxit, double eps, int checkinterval, int* iters) {
  State& s = g_st;
  if (!s.init) return GFX_ERR_STATE;
  if (checkinterval < 1) checkinterval = 1;
  const int nt = s.geo.L[3];
  const int nsh = s.geo.nh / nt;  // sites of one parity in a slice
  const bool mixed = (s.precision == GFX_PREC_MIXED);
  for (int t = 0; t < nt; t++) iters[t] = -1;
  // phase per slice: 1 sweeping in FP32, 0 sweeping in FP64, -1 stopped
  std::vector<int> phase(nt, mixed ? 1 : 0);
  std::vector<unsigned char> act[2] = {std::vector<unsigned char>(nt, mixed ? 0 : 1),
                                       std::vector<unsigned char>(nt, mixed ? 1 : 0)};
  unsigned char* dact[2] = {s.active, s.active + nt};
  const bool packed = mixed && s.packed_links && eps > PACK_EPS_MIN;
  // FP64 slices of a run on packed links are polished on fixed-point links
  // and g when their stops can be decided by the fused checks (slices of
  // whole workgroups of either size) and eps is well above the fixed-point
  // rounding bound those stops carry
  const bool fxc = packed && eps > FX_EPS_BOUND_MIN && nsh % 256 == 0;
  // every exit leaves the packed forms off
  struct PackedOff {
    State& s;
    ~PackedOff() {
      s.packed_on = false;
      s.fx_on = false;
    }
  } packed_off{s};
  s.packed_on = false;
  s.fx_on = false;
  if (mixed) {
    if (ensure_f32_links(s)) return GFX_ERR_ALLOC;
    launch_f64_to_f32(s.g64, s.g32, s.geo, 2, s.stream);
  }
  if (packed) {
    if (ensure_packed_links(s)) return GFX_ERR_ALLOC;
    tune_sweep(s, FORM_P32, 3);
    if (fxc) {
      if (ensure_fx_links(s)) return GFX_ERR_ALLOC;
      launch_f64_to_fx(s.g64, s.gfx, s.geo, 2, s.stream);
      tune_sweep(s, FORM_X64, 3);
    } else {
      tune_sweep(s, FORM_P64, 3);
    }
  } else {
    if (mixed) tune_sweep(s, FORM_F32, 3);
    tune_sweep(s, FORM_F64, 3);
  }
  s.packed_on = packed;
  s.fx_on = fxc;
  // an FP64 slice stops when its max dAdA is below eps.  A check fused
  // into the sweeps on packed or fixed-point fields decides that with the
  // rounding's worst effect on Delta added (the FP64 slices sweep in the
  // reference order), so that the bound holds on the exact links and the
  // slice's FP64 g too (a slice in the narrow band below eps stops one check
  // later); a standalone check reads the FP64 link planes.  An FP32 slice
  // never stops: below eps (or the switch) it moves to FP64, whose check
  // decides its stop.
  auto below = [&](double m, bool fused_check) {
    if (!s.packed_on || !fused_check) return m < eps;
    const double r = std::sqrt(m) + (fxc ? FX_DDELTA : PACK_DDELTA64);
    return r * r < eps;
  };
  auto upload = [&]() {
    for (int pr = 0; pr < 2; pr++)
      hipMemcpyAsync(dact[pr], act[pr].data(), nt, hipMemcpyHostToDevice, s.stream);
  };
  upload();
  s.run_iters32 = s.run_iters64 = 0;
  auto count = [&](int pr) {
    int c = 0;
    for (int t = 0; t < nt; t++) c += act[pr][t] ? 1 : 0;
    return c;
  };
  auto maps = [&]() {
    build_slice_map(s, false, act[0], nsh);
    if (mixed) build_slice_map(s, true, act[1], nsh);
  };
  maps();
  // fused checks need whole blocks per slice, in both precisions when mixed
  // (a slice moving to FP64 takes the FP32 check sweep's update with it)
  const bool fusable =
      nsh % sweep_variant_block(s.variant[sweep_form(s, false, 3)][0]) == 0 &&
      (!mixed || nsh % sweep_variant_block(s.variant[sweep_form(s, true, 3)][0]) == 0);
  auto fused_ok = [&](int) { return fusable; };
  const int stall_checks = (s.stall_iters + checkinterval - 1) / checkinterval;
  std::vector<double> best(nt, 1e300), res(4 * nt, 0.0);
  std::vector<int> stall(nt, 0);
  int left = nt;
  int even[2] = {0, 0};
  bool even_done[2] = {false, false};
  hipStreamSynchronize(s.stream);
  const auto t0 = std::chrono::steady_clock::now();
  // copy the even sites of slice t from field `from` to field `to` (planes;
  // the FP64 slices' fields are the fixed-point ones with fxc)
  auto copy_slice = [&](bool f32, int from, int to, int t) {
    const bool fx = !f32 && fxc;
    const int np = (f32 || fx) ? 3 : 6;
    const size_t vb = f32 ? sizeof(float4) : fx ? sizeof(int4) : sizeof(double2);
    char* src = f32 ? (char*)field32(s, from) : fx ? (char*)fieldfx(s, from) : (char*)field64(s, from);
    char* dst = f32 ? (char*)field32(s, to) : fx ? (char*)fieldfx(s, to) : (char*)field64(s, to);
    for (int c = 0; c < np; c++) {
      const size_t off = ((size_t)c * s.geo.stride + (size_t)t * nsh) * vb;
      hipMemcpyAsync(dst + off, src + off, (size_t)nsh * vb, hipMemcpyDeviceToDevice, s.stream);
    }
  };
  int i;
  for (i = 0; i < maxit && left > 0; i++) {
    const bool chk = (i % checkinterval) == 0;
    int n[2] = {count(0), count(1)};
    bool fused[2] = {false, false};
    for (int pr = 0; pr < 2; pr++) {
      if (!n[pr]) continue;
      fused[pr] = chk && fused_ok(pr);
      if (!even_done[pr]) do_sweep(s, pr == 1, 0, 3, dact[pr], even[pr], 0);
      do_sweep(s, pr == 1, 1, 3, dact[pr], even[pr], fused[pr] ? 1 : 0);
      even_done[pr] = false;
      if (pr == 1) s.run_iters32++; else s.run_iters64++;
    }
    if (!chk) continue;
    // the checks of state i, per precision
    for (int pr = 0; pr < 2; pr++) {
      if (!n[pr]) continue;
      if (fused[pr]) {
        do_sweep(s, pr == 1, 0, 3, dact[pr], even[pr], 2);
        const double* pc = check_partials(s, pr == 1);
        const int bps = nsh / sweep_variant_block(s.variant[sweep_form(s, pr == 1, 3)][0]);
        launch_slice_finalize(pc, pc + 3 * (size_t)s.ncheck, bps, nt, dact[pr], 3.0 * 2 * nsh,
                              2.0 * nsh, eps, s.results, s.stream);
        if (hip_check(hipMemcpyAsync(s.h_results, s.results, sizeof(double) * 4 * nt, hipMemcpyDeviceToHost, s.stream), "D2H results") ||
            hip_check(hipStreamSynchronize(s.stream), "check"))
          return GFX_ERR_DEVICE;
      } else if (functional_now(s, pr == 1, 3, nt, eps)) {
        return GFX_ERR_DEVICE;
      }
      for (int t = 0; t < nt; t++)
        if (act[pr][t])
          for (int k = 0; k < 4; k++) res[4 * t + k] = s.h_results[4 * t + k];
    }
    // decisions
    std::vector<int> stop64, sw32;
    for (int t = 0; t < nt; t++) {
      const double m = res[4 * t + 2];
      if (act[0][t]) {
        if (below(m, fused[0])) stop64.push_back(t);
      } else if (act[1][t]) {
        bool leave = m < s.switch_theta || m < eps;
        if (!leave) {
          if (m < 0.5 * best[t]) {
            best[t] = m;
            stall[t] = 0;
          } else if (best[t] < s.stall_gate && ++stall[t] >= stall_checks) {
            leave = true;
          }
        }
        if (leave) sw32.push_back(t);
      }
    }
    const bool last = i + 1 >= maxit;
    // FP64: stopped slices keep their checked state, the others take the
    // check sweep's update
    const bool swap64 = fused[0] && n[0] > 0 && !last;
    if (swap64) {
      for (int t : stop64) copy_slice(false, even[0], 2 - even[0], t);
      even[0] = 2 - even[0];
      even_done[0] = true;
    }
    // FP32 slices that leave: to FP64.  A slice going on in FP64 continues
    // from the FP32 check sweep's update when the FP64 slices also took
    // theirs (their next sweep is the odd half), else from the checked state.
    const bool swap32 = fused[1] && n[1] > 0 && !last;
    for (int t : sw32) {
      const bool from_update = swap32 && (even_done[0] || n[0] == 0);
      if (fxc)
        launch_g32_to_fx_range(field32(s, from_update ? 2 - even[1] : even[1]), field32(s, 1),
                               fieldfx(s, even[0]), fieldfx(s, 1), s.geo, t * nsh, nsh, s.stream);
      else
        launch_g32_to_g64_range(field32(s, from_update ? 2 - even[1] : even[1]), field32(s, 1),
                                field64(s, even[0]), field64(s, 1), s.geo, t * nsh, nsh, s.stream);
    }
    if (!sw32.empty() && n[0] == 0 && swap32) even_done[0] = true;
    if (swap32) {
      even[1] = 2 - even[1];
      even_done[1] = true;
    }
    for (int t : stop64) {
      act[0][t] = 0;
      phase[t] = -1;
      iters[t] = i;
      left--;
    }
    for (int t : sw32) {
      act[1][t] = 0;
      act[0][t] = 1;
      phase[t] = 0;
    }
    if (!stop64.empty() || !sw32.empty()) {
      upload();
      maps();
    }
    if (s.verbose) {
      double F = 0, D = 0, M = 0;
      for (int t = 0; t < nt; t++) {
        F += res[4 * t];
        D += res[4 * t + 1];
        M = fmax(M, res[4 * t + 2]);
      }
      printf("iter %6d:\t FUNC = %.16e \t dAdA = %.16e\t max(dAdA) = %.16e\t slices left = %d\n", i, F / nt, D / nt, M, left);
    }
  }
  // the FP64 slices' fixed-point g to FP64 (reunitarised), then slices
  // still in FP32 at maxit: their state to FP64
  if (fxc) {
    canonical_slot(s, false, even[0], 0, true);
    even[0] = 0;
    launch_fx_to_g64(s.gfx, s.g64, s.geo, s.stream);
  }
  for (int t = 0; t < nt; t++)
    if (act[1][t])
      launch_g32_to_g64_range(field32(s, even[1]), field32(s, 1), field64(s, even[0]),
                              field64(s, 1), s.geo, t * nsh, nsh, s.stream);
  canonical_even(s, false, even[0]);
  hipStreamSynchronize(s.stream);
  s.smap_n[0] = s.smap_n[1] = 0;
  s.run_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return left == 0 ? GFX_OK : GFX_ERR_NOCONV;
}

int get_trafo(gfx_su3* trafo) {
  State& s = g_st;
  if (!s.init) return GFX_ERR_STATE;
  return download_aos(s, s.g64, (double*)trafo, 1);
}
int set_trafo(gfx_su3* trafo) {
  State& s = g_st;
  if (!s.init) return GFX_ERR_STATE;
  return upload_aos(s, (const double*)trafo, s.g64, 1);
}
int get_gauge(gfx_su3* gf) {
  State& s = g_st;
  if (!s.init) return GFX_ERR_STATE;
  return download_aos(s, s.U64, (double*)gf, 4);
}

int apply_trafo(gfx_su3* gf) {
  State& s = g_st;
  if (!s.init) return GFX_ERR_STATE;
  if (ensure_staging(s)) return GFX_ERR_ALLOC;
  const int V = 2 * s.geo.nh;
  double* host = (double*)gf;
  for (int s0 = 0; s0 < V; s0 += CHUNK) {
    const int ns = (V - s0 < CHUNK) ? V - s0 : CHUNK;
    const size_t bytes = (size_t)ns * 4 * 18 * sizeof(double);
    hipMemcpyAsync(s.staging, host + (size_t)s0 * 72, bytes, hipMemcpyHostToDevice, s.stream);
    launch_apply_trafo_aos(s.staging, s.g64, s.geo, s0, ns, s.stream);
    hipMemcpyAsync(host + (size_t)s0 * 72, s.staging, bytes, hipMemcpyDeviceToHost, s.stream);
    if (hip_check(hipStreamSynchronize(s.stream), "apply_trafo")) return GFX_ERR_DEVICE;
  }
  return GFX_OK;
}

static int form_of_precision(int precision) {
  if (precision == GFX_PREC_MIXED) return FORM_F32;
  if (precision == GFX_SWEEP_PACKED64) return FORM_P64;
  if (precision == GFX_SWEEP_PACKED32) return FORM_P32;
  if (precision == GFX_SWEEP_FIXED64) return FORM_X64;
  return FORM_F64;
}

int gfx_time_sweeps(int nsweeps, int precision, double* ms) {
  State& s = g_st;
  if (!s.init) return GFX_ERR_STATE;
  const int form = form_of_precision(precision);
  const bool fp32 = form == FORM_F32 || form == FORM_P32;
  const bool fx = form == FORM_X64;
  const bool packed = form == FORM_P32 || form == FORM_P64 || fx;
  int rc = fx ? ensure_fx_links(s) : packed ? ensure_packed_links(s) : fp32 ? ensure_f32_links(s) : GFX_OK;
  if (rc) return rc;
  if (fp32) launch_f64_to_f32(s.g64, s.g32, s.geo, 2, s.stream);
  if (fx) launch_f64_to_fx(s.g64, s.gfx, s.geo, 2, s.stream);
  tune_sweep(s, form, 4);
  s.packed_on = packed;
  s.fx_on = fx;
  hipEventRecord(s.ev0, s.stream);
  for (int k = 0; k < nsweeps; k++)
    for (int p = 0; p < 2; p++) do_sweep(s, fp32, p, 4, nullptr);
  hipEventRecord(s.ev1, s.stream);
  s.packed_on = false;
  s.fx_on = false;
  if (hip_check(hipEventSynchronize(s.ev1), "timing")) return GFX_ERR_DEVICE;
  float t = 0.f;
  hipEventElapsedTime(&t, s.ev0, s.ev1);
  if (fp32) {
    launch_g32_to_g64(s.g32, s.g64, s.geo, s.stream);
    hipStreamSynchronize(s.stream);
  }
  if (fx) {
    launch_fx_to_g64(s.gfx, s.g64, s.geo, s.stream);
    hipStreamSynchronize(s.stream);
  }
  *ms = (double)t / (2.0 * nsweeps);
  return GFX_OK;
}

int gfx_set_sweep_sampling(int on) {
  State& s = g_st;
  if (on && s.sample_ev == nullptr) {
    s.sample_ev = new hipEvent_t[2 * State::SAMPLE_MAX];
    for (int k = 0; k < 2 * State::SAMPLE_MAX; k++)
      if (hip_check(hipEventCreate(&s.sample_ev[k]), "sampling events")) {
        for (int j = 0; j < k; j++) hipEventDestroy(s.sample_ev[j]);
        delete[] s.sample_ev;
        s.sample_ev = nullptr;
        return GFX_ERR_DEVICE;
      }
  }
  if (s.sample_open) close_sample(s);
  s.sample_on = on ? 1 : 0;
  s.sample_n = 0;
  return GFX_OK;
}

int gfx_sweep_sample_stats(int precision, double* mean_ms, double* min_ms, int* count) {
  State& s = g_st;
  const int form = form_of_precision(precision);
  if (s.sample_open) close_sample(s);
  if (s.sample_n > 0 && hip_check(hipEventSynchronize(s.sample_ev[2 * s.sample_n - 1]), "sampling"))
    return GFX_ERR_DEVICE;
  double sum = 0.0, mn = 0.0;
  int n = 0;
  for (int k = 0; k < s.sample_n; k++) {
    if (s.sample_form[k] != form || s.sample_cnt[k] == 0) continue;
    float t = 0.f;
    if (hip_check(hipEventElapsedTime(&t, s.sample_ev[2 * k], s.sample_ev[2 * k + 1]), "sampling"))
      return GFX_ERR_DEVICE;
    const double per = (double)t / s.sample_cnt[k];
    mn = (n == 0 || per < mn) ? per : mn;
    sum += t;
    n += s.sample_cnt[k];
  }
  if (mean_ms) *mean_ms = n ? sum / n : 0.0;
  if (min_ms) *min_ms = mn;
  if (count) *count = n;
  return GFX_OK;
}

int gfx_last_run_stats(double* ms, int* it32, int* it64) {
  if (ms) *ms = g_st.run_ms;
  if (it32) *it32 = g_st.run_iters32;
  if (it64) *it64 = g_st.run_iters64;
  return GFX_OK;
}

int gfx_sync(void) { return hip_check(hipDeviceSynchronize(), "sync"); }

// A device-to-device copy of `bytes` bytes with a plain streaming kernel,
// nrep times; returns the mean kernel time in ms.  The reference point for
// calibrating hardware byte counters (FETCH_SIZE / WRITE_SIZE).
int gfx_stream_copy(size_t bytes, int nrep, double* ms) {
  State& s = g_st;
  if (ensure_common(0)) return GFX_ERR_DEVICE;
  bytes &= ~(size_t)15;
  void *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) {
    hipFree(a);
    return GFX_ERR_ALLOC;
  }
  hipMemsetAsync(a, 0, bytes, s.stream);
  launch_stream_copy(a, b, bytes, s.stream);
  hipEventRecord(s.ev0, s.stream);
  for (int k = 0; k < nrep; k++) launch_stream_copy(a, b, bytes, s.stream);
  hipEventRecord(s.ev1, s.stream);
  int rc = hip_check(hipEventSynchronize(s.ev1), "copy");
  float t = 0.f;
  hipEventElapsedTime(&t, s.ev0, s.ev1);
  if (ms) *ms = t / (nrep > 0 ? nrep : 1);
  hipFree(a);
  hipFree(b);
  return rc;
}

int gfx_set_autotune(int on) {
  g_st.autotune = on ? 1 : 0;
  return GFX_OK;
}

int gfx_get_sweep_variant(int precision, int ndir, int* variant, int* block, int* rolled,
                          double* ms_per_variant) {
  const int pi = form_of_precision(precision), di = ndir == 4 ? 1 : 0;
  const int v = g_st.variant[pi][di];
  if (variant) *variant = v;
  if (block) *block = sweep_variant_block(v);
  if (rolled) *rolled = sweep_variant_rolled(v) ? 1 : 0;
  if (ms_per_variant)
    for (int k = 0; k < SWEEP_VARIANTS; k++) ms_per_variant[k] = g_st.tune_ms[pi][di][k];
  return g_st.tuned[pi][di] ? GFX_OK : GFX_ERR_STATE;
}

int gfx_set_sweep_variant(int precision, int ndir, int variant) {
  if (variant < 0 || variant >= SWEEP_VARIANTS) return GFX_ERR_STATE;
  const int pi = form_of_precision(precision), di = ndir == 4 ? 1 : 0;
  g_st.variant[pi][di] = variant;
  g_st.tuned[pi][di] = true;
  return GFX_OK;
}

}  // extern "C"
