// gfx_kernels.hip -- gfx950 kernels of the gauge-fixing path.
//
//   k_or_sweep       red-black overrelaxation half sweep (Landau: 4 dirs,
//                    Coulomb: 3 spatial dirs with per-slice activity),
//                    overrelax.cu:335-446
//   k_functional     per-site F and Re tr(Delta Delta^+), wavefront-shuffle
//                    + LDS block reduction, overrelax.cu:482-605
//   k_finalize       grid-level pass over block partials, overrelax.cu:610-671
//   k_plaquette      mean plaquette, overrelax.cu:34-138
//   k_pack_* / k_unpack_*   host su3 (AoS, 18 doubles) <-> device planes,
//                    su3manip.c:212-230, 281-299, 303-359
//   k_apply_trafo_aos  U_mu(x) <- g(x) U_mu(x) g^+(x+mu) on host-layout
//                    links, observables.c:17-47
//   k_f64_to_f32 / k_g32_to_g64  precision conversions of the mixed path
//
// One thread per site; all SU(3) algebra in registers (gfx_su3.h).  The
// work is HBM-bound 3x3 complex algebra, so nothing here uses MFMA.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "gfx_internal.h"
#include "gfx_lattice.h"
#include "gfx_su3.h"

#include "gfx_sweep.h"

namespace gfx {

// The check-carrying sweeps of the fixed-point polish compile to 170 VGPRs,
// two above the 168 at which a SIMD holds 3 waves: they ask for 3 waves per
// SIMD (the second bound is amdgpu_waves_per_eu), as the plain polish runs.
template <typename V, int CHK>
constexpr int sweep_min_waves() {
  return (std::is_same<V, int4>::value && CHK != 0) ? 3 : 1;
}

template <typename V, typename R, int NDIR, int BLOCK, bool ROLLED, bool NT, int CHK>
__global__ __launch_bounds__(BLOCK, (sweep_min_waves<V, CHK>())) void k_or_sweep(
    const V* __restrict__ U, V* ge, V* go, V* galt,
                                                     Geom geo, int parity, R c1, R c2,
                                                     const unsigned char* __restrict__ active,
                                                     const int* __restrict__ bmap,
                                                     double* __restrict__ partials) {
  or_sweep_block<V, R, NDIR, BLOCK, ROLLED, NT, CHK>(U, ge, go, galt, geo, parity, c1, c2, active,
                                                     bmap, partials);
}

// The half sweeps of the mixed path on packed links (PackedLinks), Landau
// (NDIR = 4) and Coulomb (NDIR = 3, active slices only):
// R = float, g in FP32 planes, links read as 21-bit fixed point (400 instead
// of 528 bytes per site update); R = double, g in FP64 planes and the
// reference's FP64 operation order, links as 24 + 16-bit fixed point within
// 2^-38 of the FP64 links (768 instead of 1056 bytes per site update).
template <typename V, typename R, int NDIR, int BLOCK, bool ROLLED, bool NT, int CHK>
__global__ __launch_bounds__(BLOCK) void k_or_sweep_packed(PackedLinks U, V* ge, V* go, V* galt,
                                                            Geom geo, int parity, R c1, R c2,
                                                            const unsigned char* __restrict__ active,
                                                            const int* __restrict__ bmap,
                                                            double* __restrict__ partials) {
  or_sweep_block<V, R, NDIR, BLOCK, ROLLED, NT, CHK>(U, ge, go, galt, geo, parity, c1, c2, active,
                                                     bmap, partials);
}

// ------------------------------------------------------------ batched sweep
template <typename V, typename R, int NDIR, int BLOCK, bool ROLLED, bool NT, int CHK, bool PK>
__global__ __launch_bounds__(BLOCK, (sweep_min_waves<V, CHK>())) void k_batch_sweep(const BatchProblem* __restrict__ probs,
                                                        const int2* __restrict__ bmap,
                                                        const int* __restrict__ active,
                                                        const int* __restrict__ poff, int parity,
                                                        R c1, R c2, double* __restrict__ partials) {
  batch_sweep_item<V, R, NDIR, BLOCK, ROLLED, NT, CHK, PK>(probs, bmap[blockIdx.x], active, poff,
                                                           parity, c1, c2, partials);
}

// grid-level pass of a fused Coulomb check, one block per active time slice
// t over its bps block partials [t bps, (t+1) bps) of each sweep:
// out[4t..] = F/normF, dAdA/normD, max dAdA, max < eps.  Inactive slices
// keep their last values.
__global__ __launch_bounds__(BLK) void k_slice_finalize(const double* __restrict__ podd,
                                                        const double* __restrict__ peven, int bps,
                                                        const unsigned char* __restrict__ active,
                                                        double normF, double normD, double eps,
                                                        double* __restrict__ out) {
  const int t = blockIdx.x;
  if (!active[t]) return;
  double F = 0.0, d = 0.0, mx = 0.0;
#pragma unroll 8
  for (int k = threadIdx.x; k < bps; k += BLK) {
    const size_t b = (size_t)t * bps + k;
    F += podd[3 * b];
    d += podd[3 * b + 1] + peven[3 * b + 1];
    mx = fmax(mx, fmax(podd[3 * b + 2], peven[3 * b + 2]));
  }
  block_reduce3(F, d, mx);
  if (threadIdx.x == 0) {
    out[4 * t + 0] = F / normF;
    out[4 * t + 1] = d / normD;
    out[4 * t + 2] = mx;
    out[4 * t + 3] = (mx < eps) ? 1.0 : 0.0;
  }
}

// grid-level pass of a fused check: podd / peven hold nb block partials
// each.  out = F/normF (F from the odd sweep), dAdA/normD, max dAdA, and
// 1.0 if max < eps.
__global__ __launch_bounds__(BLK) void k_check_finalize(const double* __restrict__ podd,
                                                        const double* __restrict__ peven, int nb,
                                                        double normF, double normD, double eps,
                                                        double* __restrict__ out) {
  double F = 0.0, d = 0.0, mx = 0.0;
  // unrolled so that the loads of several blocks are in flight together (the
  // sums keep their order)
#pragma unroll 8
  for (int b = threadIdx.x; b < nb; b += BLK) {
    F += podd[3 * (size_t)b];
    d += podd[3 * (size_t)b + 1] + peven[3 * (size_t)b + 1];
    mx = fmax(mx, fmax(podd[3 * (size_t)b + 2], peven[3 * (size_t)b + 2]));
  }
  block_reduce3(F, d, mx);
  if (threadIdx.x == 0) {
    out[0] = F / normF;
    out[1] = d / normD;
    out[2] = mx;
    out[3] = (mx < eps) ? 1.0 : 0.0;
  }
}

// ------------------------------------------------------------ functional
// Segment = whole lattice (Landau) or one time slice (Coulomb).  Within a
// segment, sites [0, nsh) are parity 0 and [nsh, 2 nsh) parity 1; the
// parity index of site k of segment seg is seg*nsh + (k mod nsh).
// F contribution (sum over mu of Re tr g(x) U_mu(x) g^+(x+mu)) and
// Re tr(Delta Delta^+) of site i of parity p (overrelax.cu:498-586), in FP64
// whatever the storage precision.  The direction loop is kept rolled.
template <typename V, int NDIR>
__device__ __forceinline__ void functional_site(const V* __restrict__ U, const V* __restrict__ ge,
                                                const V* __restrict__ go, const Geom& geo, int p,
                                                int i, double& F, double& d) {
  const int S = geo.stride;
  const Site s = site_coords(geo, i, p);
  const V* gme = p ? go : ge;
  const V* gop = p ? ge : go;
  int jf[NDIR], jb[NDIR];
#pragma unroll
  for (int mu = 0; mu < NDIR; mu++) {
    jf[mu] = nb_fwd(geo, s, i, mu);
    jb[mu] = nb_bwd(geo, s, i, mu);
  }
  M3<double> gather;
  m_zero(gather);
  F = 0.0;
#pragma unroll 1
  for (int mu = 0; mu < NDIR; mu++) {
    const M3<double> Gx = load_m<V, double>(gme, S, i);
    M3<double> Um = load_m<V, double>(field_ptr(U, p * 4 + mu, S), S, i);
    M3<double> help = m_mul(Gx, Um);
    M3<double> Gd = load_m_dagger<V, double>(gop, S, jf[mu]);
    M3<double> fw = m_mul(help, Gd);
    F += m_retrace(fw);
    Um = load_m<V, double>(field_ptr(U, (1 - p) * 4 + mu, S), S, jb[mu]);
    M3<double> Gb = load_m<V, double>(gop, S, jb[mu]);
    help = m_mul(Gb, Um);
    const M3<double> Gxd = load_m_dagger<V, double>(gme, S, i);
    M3<double> bw = m_mul(help, Gxd);
    vecpot(fw);
    vecpot(bw);
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
      for (int c = 0; c < 3; c++) {
        Cx<double> t = csub(bw.e[a][c], fw.e[a][c]);
        gather.e[a][c] = cadd(gather.e[a][c], t);
      }
  }
  M3<double> dd = m_mul_d(gather, gather);
  d = m_retrace(dd);
}

template <typename V, int NDIR>
__global__ __launch_bounds__(BLK) void k_functional(const V* __restrict__ U,
                                                     const V* __restrict__ ge,
                                                     const V* __restrict__ go, Geom geo, int nsh,
                                                     double* __restrict__ partials) {
  const int seg = blockIdx.y;
  const int k = blockIdx.x * BLK + threadIdx.x;
  double F = 0.0, d = 0.0, mx = 0.0;
  if (k < 2 * nsh) {
    const int p = (k >= nsh) ? 1 : 0;
    const int i = seg * nsh + (k - p * nsh);
    functional_site<V, NDIR>(U, ge, go, geo, p, i, F, d);
    mx = d;
  }
  block_reduce3(F, d, mx);
  if (threadIdx.x == 0) {
    double* o = partials + 3 * ((size_t)seg * gridDim.x + blockIdx.x);
    o[0] = F;
    o[1] = d;
    o[2] = mx;
  }
}

// Grid-level pass of a batch's fused check, one block per lattice.  A
// lattice is in the FP64 phase (act64), the FP32 phase (act32) or stopped;
// its block partials of the odd (pod) and even (pev) check sweeps of its
// phase lie at [poff[q], poff[q+1]) of that phase's arrays; inmap64 /
// inmap32 say whose blocks the sweeps of this period launched.
// results[4q..] = F/(ndir V), dAdA/V, max dAdA, stop iteration.
// FP64: stop when max < eps, else take the check sweep's even field (swap
// g[0] and galt) unless *iter + 1 == maxit (the run ends on the checked
// state).  FP32: leave the phase when max < eps or max < switch_theta or
// when the best max has stayed below stall_gate without halving for
// stall_checks checks; the lattice then continues in FP64 from the FP32
// state after the check sweep (conv = 1: even field from galt), or, when the
// run ends, keeps the checked state converted (conv = 2, not stopped).
// sw[q] records the iteration of the switch.
__global__ __launch_bounds__(BLK) void k_batch_finalize(BatchCheck c, int ndir, double eps,
                                                       double switch_theta, int maxit,
                                                       int stall_checks, double stall_gate) {
  batch_finalize_lattice(blockIdx.x, c, ndir, eps, switch_theta, maxit, stall_checks, stall_gate,
                         nullptr);
}

// FP32 -> FP64 g of the lattices with conv[q] != 0 (FP64 Gram-Schmidt
// reunitarisation, as k_g32_to_g64), stored as FP64 (VD = double2) or as
// 32-bit fixed point (VD = int4): bmap of BLK-site blocks, y = parity.
// conv 1 takes the even field from galt (the check sweep's update), conv 2
// from g[0] (the checked state).
template <typename VD>
__global__ __launch_bounds__(BLK) void k_batch_convert(const BatchProblem* __restrict__ probs32,
                                                       const BatchProblem* __restrict__ probs64,
                                                       const int2* __restrict__ bmap,
                                                       const int* __restrict__ conv) {
  const int2 m = bmap[blockIdx.x];
  const int cv = conv[m.x];
  if (!cv) return;
  const int parity = blockIdx.y;
  const BatchProblem& P = probs32[m.x];
  const int i = m.y * BLK + threadIdx.x;
  if (i >= P.geo.nh) return;
  const float4* src = static_cast<const float4*>(parity ? P.g[1] : (cv == 1 ? P.galt : P.g[0]));
  VD* dst = static_cast<VD*>(probs64[m.x].g[parity]);
  M3<double> M = load_m<float4, double>(src, P.geo.stride, i);
  m_normalize(M);
  Store<VD>::store(dst, P.geo.stride, i, M);
}

__global__ void k_batch_advance(int* iter, int step) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *iter += step;
}

// fills used instead of memsets (values travel as kernel arguments)
__global__ void k_fill_i32(int* __restrict__ p, int n, int v) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) p[k] = v;
}
__global__ void k_fill_f64(double* __restrict__ p, int n, double v) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) p[k] = v;
}

// grid-level pass: one block per segment over its nb block partials.
// out[4*seg + {0,1,2,3}] = F/(ndir*Vseg), dAdA/Vseg, max dAdA, converged flag
__global__ __launch_bounds__(BLK) void k_finalize(const double* __restrict__ partials, int nb,
                                                  double normF, double normD, double eps,
                                                  double* __restrict__ out) {
  const int seg = blockIdx.x;
  double F = 0.0, d = 0.0, mx = 0.0;
  for (int b = threadIdx.x; b < nb; b += BLK) {
    const double* p = partials + 3 * ((size_t)seg * nb + b);
    F += p[0];
    d += p[1];
    mx = fmax(mx, p[2]);
  }
  block_reduce3(F, d, mx);
  if (threadIdx.x == 0) {
    out[4 * seg + 0] = F / normF;
    out[4 * seg + 1] = d / normD;
    out[4 * seg + 2] = mx;
    out[4 * seg + 3] = (mx < eps) ? 1.0 : 0.0;
  }
}

// ------------------------------------------------------------ plaquette
template <typename V>
__global__ __launch_bounds__(BLK) void k_plaquette(const V* __restrict__ U, Geom geo,
                                                   double* __restrict__ partials) {
  const int k = blockIdx.x * BLK + threadIdx.x;
  double acc = 0.0, dummy = 0.0, dummy2 = 0.0;
  if (k < 2 * geo.nh) {
    const int p = (k >= geo.nh) ? 1 : 0;
    const int i = k - p * geo.nh;
    const int S = geo.stride;
    const Site s = site_coords(geo, i, p);
    // loops kept rolled: unrolled, the six plaquettes' loads spill
#pragma unroll 1
    for (int nu = 0; nu < 3; nu++) {
      const int jnu = nb_fwd(geo, s, i, nu);
#pragma unroll 1
      for (int mu = nu + 1; mu < 4; mu++) {
        const int jmu = nb_fwd(geo, s, i, mu);
        M3<double> M1 = load_m<V, double>(field_ptr(U, p * 4 + mu, S), S, i);
        M3<double> M2 = load_m<V, double>(field_ptr(U, (1 - p) * 4 + nu, S), S, jmu);
        M3<double> M3d = load_m_dagger<V, double>(field_ptr(U, (1 - p) * 4 + mu, S), S, jnu);
        M3<double> M4d = load_m_dagger<V, double>(field_ptr(U, p * 4 + nu, S), S, i);
        M3<double> s1 = m_mul(M3d, M4d);
        M3<double> s2 = m_mul(M2, s1);
        s1 = m_mul(M1, s2);
        acc += m_retrace(s1) / 3.0;
      }
    }
  }
  block_reduce3(acc, dummy, dummy2);
  if (threadIdx.x == 0) partials[3 * blockIdx.x] = acc;
}

__global__ __launch_bounds__(BLK) void k_sum_partials(const double* __restrict__ partials, int nb,
                                                      double norm, double* __restrict__ out) {
  double a = 0.0, b = 0.0, c = 0.0;
  for (int k = threadIdx.x; k < nb; k += BLK) a += partials[3 * k];
  block_reduce3(a, b, c);
  if (threadIdx.x == 0) out[0] = a / norm;
}

// ------------------------------------------------------------ layout conversion
// parity and parity index of lexicographic site s
__device__ __forceinline__ void lex_site(const Geom& geo, FastDiv dL0, int s, int x[4], int& p) {
  uint32_t q = fdiv((uint32_t)s, dL0);
  x[0] = s - (int)q * geo.L[0];
  uint32_t q2 = fdiv(q, geo.dL1);
  x[1] = (int)(q - q2 * (uint32_t)geo.L[1]);
  uint32_t q3 = fdiv(q2, geo.dL2);
  x[2] = (int)(q2 - q3 * (uint32_t)geo.L[2]);
  x[3] = (int)q3;
  p = (x[0] + x[1] + x[2] + x[3]) & 1;
}

// nmat matrices per site (4 for links, 1 for g); staging holds sites
// [s0, s0+ns) as AoS su3 (18 doubles each, matrix m of site s at
// (s-s0)*nmat+m).
template <typename V>
__global__ __launch_bounds__(BLK) void k_pack(const double* __restrict__ aos, V* __restrict__ dst,
                                              Geom geo, FastDiv dL0, int s0, int ns, int nmat) {
  const int t = blockIdx.x * BLK + threadIdx.x;
  if (t >= ns * nmat) return;
  const int sl = t / nmat, m = t - sl * nmat;
  const int s = s0 + sl;
  int x[4], p;
  lex_site(geo, dL0, s, x, p);
  const double* a = aos + (size_t)t * 18;
  M3<double> M;
#pragma unroll
  for (int r = 0; r < 2; r++)
#pragma unroll
    for (int c = 0; c < 3; c++) M.e[r][c] = cx<double>(a[6 * r + 2 * c], a[6 * r + 2 * c + 1]);
  Store<V>::store(field_ptr(dst, p * nmat + m, geo.stride), geo.stride, s >> 1, M);
}

template <typename V>
__global__ __launch_bounds__(BLK) void k_unpack(const V* __restrict__ src, double* __restrict__ aos,
                                                Geom geo, FastDiv dL0, int s0, int ns, int nmat) {
  const int t = blockIdx.x * BLK + threadIdx.x;
  if (t >= ns * nmat) return;
  const int sl = t / nmat, m = t - sl * nmat;
  const int s = s0 + sl;
  int x[4], p;
  lex_site(geo, dL0, s, x, p);
  M3<double> M;
  Store<V>::template load<double>(field_ptr(src, p * nmat + m, geo.stride), geo.stride, s >> 1,
                                  M.e[0], M.e[1]);
  m_row2_unfused(M);
  double* a = aos + (size_t)t * 18;
#pragma unroll
  for (int r = 0; r < 3; r++)
#pragma unroll
    for (int c = 0; c < 3; c++) {
      a[6 * r + 2 * c] = M.e[r][c].re;
      a[6 * r + 2 * c + 1] = M.e[r][c].im;
    }
}

// host-layout links of sites [s0, s0+ns) in place: U <- g(x) U g^+(x+mu)
__global__ __launch_bounds__(BLK) void k_apply_trafo_aos(double* __restrict__ aos,
                                                         const double2* __restrict__ g, Geom geo,
                                                         FastDiv dL0, int s0, int ns) {
  const int t = blockIdx.x * BLK + threadIdx.x;
  if (t >= ns * 4) return;
  const int sl = t >> 2, mu = t & 3;
  const int s = s0 + sl;
  int x[4], p;
  lex_site(geo, dL0, s, x, p);
  const int i = s >> 1;
  const Site st = site_coords(geo, i, p);
  const int jf = nb_fwd(geo, st, i, mu);
  double* a = aos + (size_t)t * 18;
  M3<double> Um;
#pragma unroll
  for (int r = 0; r < 3; r++)
#pragma unroll
    for (int c = 0; c < 3; c++) Um.e[r][c] = cx<double>(a[6 * r + 2 * c], a[6 * r + 2 * c + 1]);
  const M3<double> G = load_m<double2, double>(field_ptr(g, p, geo.stride), geo.stride, i);
  const M3<double> Gpd = load_m_dagger<double2, double>(field_ptr(g, 1 - p, geo.stride), geo.stride, jf);
  M3<double> h1 = m_mul(G, Um);
  M3<double> r = m_mul(h1, Gpd);
#pragma unroll
  for (int rr = 0; rr < 3; rr++)
#pragma unroll
    for (int c = 0; c < 3; c++) {
      a[6 * rr + 2 * c] = r.e[rr][c].re;
      a[6 * rr + 2 * c + 1] = r.e[rr][c].im;
    }
}

// nfield FP64 fields -> FP32 fields (12-real each)
__global__ __launch_bounds__(BLK) void k_f64_to_f32(const double2* __restrict__ src,
                                                    float4* __restrict__ dst, int nh, int stride,
                                                    int nfield) {
  const int i = blockIdx.x * BLK + threadIdx.x;
  const int f = blockIdx.y;
  if (i >= nh || f >= nfield) return;
  const M3<double> M = load_m<double2, double>(field_ptr(src, f, stride), stride, i);
  Store<float4>::store(field_ptr(dst, f, stride), stride, i, M);
}

// nfield FP64 link fields -> packed links (PackedLinks): 24-bit hi parts in
// the a/b/c planes, 16-bit lo parts in the lo planes, 21-bit values in the
// a21/b21 planes
__global__ __launch_bounds__(BLK) void k_pack_links(const double2* __restrict__ src,
                                                    uint4* __restrict__ pa, uint4* __restrict__ pb,
                                                    uint32_t* __restrict__ pc,
                                                    uint2* __restrict__ lo,
                                                    uint4* __restrict__ pa21,
                                                    uint4* __restrict__ pb21, int nh, int stride,
                                                    int nfield) {
  const int i = blockIdx.x * BLK + threadIdx.x;
  const int f = blockIdx.y;
  if (i >= nh || f >= nfield) return;
  const double2* u = field_ptr(src, f, stride);
  double r[12];
#pragma unroll
  for (int c = 0; c < 6; c++) {
    const double2 v = u[(size_t)c * stride + i];
    r[2 * c] = v.x;
    r[2 * c + 1] = v.y;
  }
  uint32_t w[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  uint32_t x[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  int q[12];
#pragma unroll
  for (int k = 0; k < 12; k++) {
    const int h = pack_hi_of(r[k]);
    q[k] = pack_lo_of(r[k], h);
    const uint32_t hv = (uint32_t)h & 0xffffffu;
    const int o = 24 * k, d = o / 32, sh = o % 32;
    w[d] |= hv << sh;
    if (sh > 8) w[d + 1] |= hv >> (32 - sh);
    const uint32_t v21 = (uint32_t)pack21_of(r[k]) & 0x1fffffu;
    const int o2 = 21 * k, d2 = o2 / 32, sh2 = o2 % 32;
    x[d2] |= v21 << sh2;
    if (sh2 > 11) x[d2 + 1] |= v21 >> (32 - sh2);
  }
  const size_t fi = (size_t)f * stride + i;
  pa[fi] = make_uint4(w[0], w[1], w[2], w[3]);
  pb[fi] = make_uint4(w[4], w[5], w[6], w[7]);
  pc[fi] = w[8];
  pa21[fi] = make_uint4(x[0], x[1], x[2], x[3]);
  pb21[fi] = make_uint4(x[4], x[5], x[6], x[7]);
  uint2* l = lo + (size_t)f * 3 * stride + i;
#pragma unroll
  for (int p = 0; p < 3; p++)
    l[(size_t)p * stride] = make_uint2(((uint32_t)q[4 * p] & 0xffffu) | ((uint32_t)q[4 * p + 1] << 16),
                                       ((uint32_t)q[4 * p + 2] & 0xffffu) | ((uint32_t)q[4 * p + 3] << 16));
}

// FP32 g -> FP64 g with FP64 reunitarisation (Gram-Schmidt, dev_su3.cu:1569-1634)
__global__ __launch_bounds__(BLK) void k_g32_to_g64(const float4* __restrict__ src,
                                                    double2* __restrict__ dst, int nh, int stride) {
  const int i = blockIdx.x * BLK + threadIdx.x;
  const int f = blockIdx.y;
  if (i >= nh) return;
  M3<double> M = load_m<float4, double>(field_ptr(src, f, stride), stride, i);
  m_normalize(M);
  Store<double2>::store(field_ptr(dst, f, stride), stride, i, M);
}

// the same for the sites [i0, i0 + n) of each parity, with the even and odd
// fields given separately (Coulomb: one time slice leaving FP32)
__global__ __launch_bounds__(BLK) void k_g32_to_g64_range(const float4* __restrict__ se,
                                                          const float4* __restrict__ so,
                                                          double2* __restrict__ de,
                                                          double2* __restrict__ dd, int i0, int n,
                                                          int stride) {
  const int k = blockIdx.x * BLK + threadIdx.x;
  if (k >= n) return;
  const int i = i0 + k;
  const float4* src = blockIdx.y ? so : se;
  double2* dst = blockIdx.y ? dd : de;
  M3<double> M = load_m<float4, double>(src, stride, i);
  m_normalize(M);
  Store<double2>::store(dst, stride, i, M);
}

// nfield fields between storage forms, matrix by matrix (grid.y = field):
// FP64 -> 32-bit fixed point (links and g of the fixed-point polish), and
// back with an FP64 Gram-Schmidt reunitarisation; FP32 g -> fixed point with
// the same reunitarisation (the FP32 phase handing over to the polish)
template <typename VS, typename VD, bool NORM>
__global__ __launch_bounds__(BLK) void k_convert(const VS* __restrict__ src, VD* __restrict__ dst,
                                                 int nh, int stride) {
  const int i = blockIdx.x * BLK + threadIdx.x;
  const int f = blockIdx.y;
  if (i >= nh) return;
  M3<double> M = load_m<VS, double>(field_ptr(src, f, stride), stride, i);
  if (NORM) m_normalize(M);
  Store<VD>::store(field_ptr(dst, f, stride), stride, i, M);
}

// the sites [i0, i0 + n) of each parity (one Coulomb time slice): FP32 g ->
// fixed point with the FP64 reunitarisation (grid.y = parity, fields given
// separately); and a plain copy of those sites between two fixed-point
// fields
__global__ __launch_bounds__(BLK) void k_g32_to_fx_range(const float4* __restrict__ se,
                                                         const float4* __restrict__ so,
                                                         int4* __restrict__ de,
                                                         int4* __restrict__ dd, int i0, int n,
                                                         int stride) {
  const int k = blockIdx.x * BLK + threadIdx.x;
  if (k >= n) return;
  const int i = i0 + k;
  M3<double> M = load_m<float4, double>(blockIdx.y ? so : se, stride, i);
  m_normalize(M);
  Store<int4>::store(blockIdx.y ? dd : de, stride, i, M);
}

// ------------------------------------------------------------ launchers
static inline dim3 grid1(long n) { return dim3((unsigned)((n + BLK - 1) / BLK)); }

// sweep variant v: block 64 if bit 0 else 256; unrolled star loop if bit 1;
// non-temporal link loads if bit 2
int sweep_variant_block(int v) { return (v & 1) ? 64 : 256; }
bool sweep_variant_rolled(int v) { return (v & 2) == 0; }
bool sweep_variant_nt(int v) { return (v & 4) != 0; }

template <typename V>
struct SweepArgs {
  const V* U;
  V* ge;
  V* go;
  V* galt;
  Geom geo;
  int parity;
  const unsigned char* active;
  const int* bmap;
  int nblocks;  // entries of bmap to launch (0: every block of the lattice)
  double* partials;
};

template <typename V, typename R, int NDIR, int BLOCK, bool ROLLED, bool NT, int CHK>
static void launch_sweep_v(const SweepArgs<V>& a, R c1, R c2, hipStream_t st) {
  dim3 grid(a.nblocks > 0 ? a.nblocks : (a.geo.nh + BLOCK - 1) / BLOCK);
  hipLaunchKernelGGL((k_or_sweep<V, R, NDIR, BLOCK, ROLLED, NT, CHK>), grid, dim3(BLOCK), 0, st,
                     a.U, a.ge, a.go, a.galt, a.geo, a.parity, c1, c2, a.active, a.bmap,
                     a.partials);
}

template <typename V, typename R, int NDIR, int CHK>
static void launch_sweep_chk(const SweepArgs<V>& a, R c1, R c2, int variant, hipStream_t st) {
  switch (variant & 7) {
    case 0: launch_sweep_v<V, R, NDIR, 256, true, false, CHK>(a, c1, c2, st); break;
    case 1: launch_sweep_v<V, R, NDIR, 64, true, false, CHK>(a, c1, c2, st); break;
    case 2: launch_sweep_v<V, R, NDIR, 256, false, false, CHK>(a, c1, c2, st); break;
    case 3: launch_sweep_v<V, R, NDIR, 64, false, false, CHK>(a, c1, c2, st); break;
    case 4: launch_sweep_v<V, R, NDIR, 256, true, true, CHK>(a, c1, c2, st); break;
    case 5: launch_sweep_v<V, R, NDIR, 64, true, true, CHK>(a, c1, c2, st); break;
    case 6: launch_sweep_v<V, R, NDIR, 256, false, true, CHK>(a, c1, c2, st); break;
    default: launch_sweep_v<V, R, NDIR, 64, false, true, CHK>(a, c1, c2, st); break;
  }
}

template <typename V, typename R>
static void launch_sweep_t(const SweepArgs<V>& a, int ndir, double omega, int variant, int chk,
                           hipStream_t st) {
  // C(w,1) = (1+w-1)/1 and C(w,2) = (1+w-2)/2 as the reference forms them
  // (overrelax.cu:167), in double, then rounded to R.
  const R c1 = (R)((1.0 + omega - 1) / 1);
  const R c2 = (R)((1.0 + omega - 2) / 2);
  if (ndir == 4) {
    SweepArgs<V> b = a;
    b.active = nullptr;
    if (chk == 1)
      launch_sweep_chk<V, R, 4, 1>(b, c1, c2, variant, st);
    else if (chk == 2)
      launch_sweep_chk<V, R, 4, 2>(b, c1, c2, variant, st);
    else
      launch_sweep_chk<V, R, 4, 0>(b, c1, c2, variant, st);
  } else {
    if (chk == 1)
      launch_sweep_chk<V, R, 3, 1>(a, c1, c2, variant, st);
    else if (chk == 2)
      launch_sweep_chk<V, R, 3, 2>(a, c1, c2, variant, st);
    else
      launch_sweep_chk<V, R, 3, 0>(a, c1, c2, variant, st);
  }
}

void launch_sweep_f64(const double2* U, double2* ge, double2* go, const Geom& geo, int parity,
                      int ndir, double omega, const unsigned char* active, int variant,
                      const int* bmap, hipStream_t st, int chk, double2* galt, double* partials,
                      int nblocks) {
  launch_sweep_t<double2, double>({U, ge, go, galt, geo, parity, active, bmap, nblocks, partials},
                                  ndir, omega, variant, chk, st);
}
void launch_sweep_f32(const float4* U, float4* ge, float4* go, const Geom& geo, int parity,
                      int ndir, double omega, const unsigned char* active, int variant,
                      const int* bmap, hipStream_t st, int chk, float4* galt, double* partials,
                      int nblocks) {
  launch_sweep_t<float4, float>({U, ge, go, galt, geo, parity, active, bmap, nblocks, partials},
                                ndir, omega, variant, chk, st);
}

template <typename V, typename R, int NDIR, int BLOCK, bool ROLLED, bool NT, int CHK>
static void launch_packed_v(const PackedArgs& p, const SweepArgs<V>& a, R c1, R c2,
                            hipStream_t st) {
  dim3 grid(a.nblocks > 0 ? a.nblocks : (a.geo.nh + BLOCK - 1) / BLOCK);
  const PackedLinks U{p.a, p.b, p.c, p.lo, p.a21, p.b21};
  hipLaunchKernelGGL((k_or_sweep_packed<V, R, NDIR, BLOCK, ROLLED, NT, CHK>), grid, dim3(BLOCK), 0,
                     st, U, a.ge, a.go, a.galt, a.geo, a.parity, c1, c2, a.active, a.bmap,
                     a.partials);
}

// FP64 on packed links: rolled variants only (bit 1 of the variant ignored)
template <typename V, typename R, int NDIR, int CHK>
static void launch_packed_chk(const PackedArgs& p, const SweepArgs<V>& a, R c1, R c2, int variant,
                              hipStream_t st) {
  const int v = std::is_same<R, double>::value ? (variant & ~2) : variant;
  switch (v & 7) {
    case 0: launch_packed_v<V, R, NDIR, 256, true, false, CHK>(p, a, c1, c2, st); break;
    case 1: launch_packed_v<V, R, NDIR, 64, true, false, CHK>(p, a, c1, c2, st); break;
    case 4: launch_packed_v<V, R, NDIR, 256, true, true, CHK>(p, a, c1, c2, st); break;
    case 5: launch_packed_v<V, R, NDIR, 64, true, true, CHK>(p, a, c1, c2, st); break;
    default:
      if constexpr (std::is_same<R, float>::value) {
        switch (v & 7) {
          case 2: launch_packed_v<V, R, NDIR, 256, false, false, CHK>(p, a, c1, c2, st); break;
          case 3: launch_packed_v<V, R, NDIR, 64, false, false, CHK>(p, a, c1, c2, st); break;
          case 6: launch_packed_v<V, R, NDIR, 256, false, true, CHK>(p, a, c1, c2, st); break;
          default: launch_packed_v<V, R, NDIR, 64, false, true, CHK>(p, a, c1, c2, st); break;
        }
      }
      break;
  }
}

template <typename V, typename R>
static void launch_packed_t(const PackedArgs& p, const SweepArgs<V>& a, int ndir, double omega,
                            int variant, int chk, hipStream_t st) {
  const R c1 = (R)((1.0 + omega - 1) / 1);
  const R c2 = (R)((1.0 + omega - 2) / 2);
  if (ndir == 4) {
    SweepArgs<V> b = a;
    b.active = nullptr;
    if (chk == 1)
      launch_packed_chk<V, R, 4, 1>(p, b, c1, c2, variant, st);
    else if (chk == 2)
      launch_packed_chk<V, R, 4, 2>(p, b, c1, c2, variant, st);
    else
      launch_packed_chk<V, R, 4, 0>(p, b, c1, c2, variant, st);
  } else {
    if (chk == 1)
      launch_packed_chk<V, R, 3, 1>(p, a, c1, c2, variant, st);
    else if (chk == 2)
      launch_packed_chk<V, R, 3, 2>(p, a, c1, c2, variant, st);
    else
      launch_packed_chk<V, R, 3, 0>(p, a, c1, c2, variant, st);
  }
}

void launch_sweep_packed_f32(const PackedArgs& p, float4* ge, float4* go, const Geom& geo,
                             int parity, int ndir, double omega, const unsigned char* active,
                             int variant, const int* bmap, hipStream_t st, int chk, float4* galt,
                             double* partials, int nblocks) {
  launch_packed_t<float4, float>(p, {nullptr, ge, go, galt, geo, parity, active, bmap, nblocks, partials},
                                 ndir, omega, variant, chk, st);
}
void launch_sweep_packed_f64(const PackedArgs& p, double2* ge, double2* go, const Geom& geo,
                             int parity, int ndir, double omega, const unsigned char* active,
                             int variant, const int* bmap, hipStream_t st, int chk, double2* galt,
                             double* partials, int nblocks) {
  launch_packed_t<double2, double>(p, {nullptr, ge, go, galt, geo, parity, active, bmap, nblocks, partials},
                                   ndir, omega, variant, chk, st);
}

// the FP64 polish on 32-bit fixed-point links and g (Landau: 4 directions;
// Coulomb: 3, active slices only): bit 1 of the variant selects the star one
// term ahead (the two operands of the next term load while a term is formed,
// 164-166 VGPRs) instead of the rolled direction loop (the fully unrolled
// FP64 star would spill)
template <int NDIR, int CHK>
static void launch_fx_chk(const SweepArgs<int4>& a, double c1, double c2, int variant,
                          hipStream_t st) {
  switch (variant & 7) {
    case 0: launch_sweep_v<int4, double, NDIR, 256, true, false, CHK>(a, c1, c2, st); break;
    case 1: launch_sweep_v<int4, double, NDIR, 64, true, false, CHK>(a, c1, c2, st); break;
    case 2: launch_sweep_v<int4, double, NDIR, 256, false, false, CHK>(a, c1, c2, st); break;
    case 3: launch_sweep_v<int4, double, NDIR, 64, false, false, CHK>(a, c1, c2, st); break;
    case 4: launch_sweep_v<int4, double, NDIR, 256, true, true, CHK>(a, c1, c2, st); break;
    case 5: launch_sweep_v<int4, double, NDIR, 64, true, true, CHK>(a, c1, c2, st); break;
    case 6: launch_sweep_v<int4, double, NDIR, 256, false, true, CHK>(a, c1, c2, st); break;
    default: launch_sweep_v<int4, double, NDIR, 64, false, true, CHK>(a, c1, c2, st); break;
  }
}
template <int NDIR>
static void launch_fx_ndir(const SweepArgs<int4>& a, double c1, double c2, int variant, int chk,
                           hipStream_t st) {
  if (chk == 1)
    launch_fx_chk<NDIR, 1>(a, c1, c2, variant, st);
  else if (chk == 2)
    launch_fx_chk<NDIR, 2>(a, c1, c2, variant, st);
  else
    launch_fx_chk<NDIR, 0>(a, c1, c2, variant, st);
}
void launch_sweep_fx(const int4* U, int4* ge, int4* go, const Geom& geo, int parity, int ndir,
                     double omega, const unsigned char* active, int variant, const int* bmap,
                     hipStream_t st, int chk, int4* galt, double* partials, int nblocks) {
  const double c1 = (1.0 + omega - 1) / 1;
  const double c2 = (1.0 + omega - 2) / 2;
  if (ndir == 4) {
    const SweepArgs<int4> a{U, ge, go, galt, geo, parity, nullptr, bmap, nblocks, partials};
    launch_fx_ndir<4>(a, c1, c2, variant, chk, st);
  } else {
    const SweepArgs<int4> a{U, ge, go, galt, geo, parity, active, bmap, nblocks, partials};
    launch_fx_ndir<3>(a, c1, c2, variant, chk, st);
  }
}

int sweep_blocks(int nh, int variant) {
  const int b = sweep_variant_block(variant);
  return (nh + b - 1) / b;
}

void launch_slice_finalize(const double* podd, const double* peven, int bps, int nt,
                           const unsigned char* active, double normF, double normD, double eps,
                           double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_slice_finalize, dim3(nt), dim3(BLK), 0, st, podd, peven, bps, active, normF,
                     normD, eps, out);
}

void launch_check_finalize(const double* podd, const double* peven, int nb, double normF,
                           double normD, double eps, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_check_finalize, dim3(1), dim3(BLK), 0, st, podd, peven, nb, normF, normD,
                     eps, out);
}

// ---- batch launchers (Landau): nblocks entries of bmap built for the
// variant's block size, poff the partial offsets of that map
struct BSweepArgs {
  const BatchProblem* probs;
  const int2* bmap;
  int nblocks;
  const int* active;
  const int* poff;
  int parity;
  double* partials;
  bool packed;
};

template <typename V, typename R, int BLOCK, bool ROLLED, bool NT, int CHK>
static void launch_bsweep_v(const BSweepArgs& a, R c1, R c2, hipStream_t st) {
  if constexpr (std::is_same<V, int4>::value) {
    hipLaunchKernelGGL((k_batch_sweep<V, R, 4, BLOCK, ROLLED, NT, CHK, false>), dim3(a.nblocks),
                       dim3(BLOCK), 0, st, a.probs, a.bmap, a.active, a.poff, a.parity, c1, c2,
                       a.partials);
  } else if (a.packed) {
    hipLaunchKernelGGL((k_batch_sweep<V, R, 4, BLOCK, ROLLED, NT, CHK, true>), dim3(a.nblocks),
                       dim3(BLOCK), 0, st, a.probs, a.bmap, a.active, a.poff, a.parity, c1, c2,
                       a.partials);
  } else {
    hipLaunchKernelGGL((k_batch_sweep<V, R, 4, BLOCK, ROLLED, NT, CHK, false>), dim3(a.nblocks),
                       dim3(BLOCK), 0, st, a.probs, a.bmap, a.active, a.poff, a.parity, c1, c2,
                       a.partials);
  }
}

template <typename V, typename R, int CHK>
static void launch_bsweep_chk(const BSweepArgs& a, R c1, R c2, int variant, hipStream_t st) {
  // on fixed-point fields bit 1 selects the star one term ahead instead of
  // the rolled direction loop, as in the single-lattice polish
  switch (variant & 7) {
    case 0: launch_bsweep_v<V, R, 256, true, false, CHK>(a, c1, c2, st); break;
    case 1: launch_bsweep_v<V, R, 64, true, false, CHK>(a, c1, c2, st); break;
    case 2: launch_bsweep_v<V, R, 256, false, false, CHK>(a, c1, c2, st); break;
    case 3: launch_bsweep_v<V, R, 64, false, false, CHK>(a, c1, c2, st); break;
    case 4: launch_bsweep_v<V, R, 256, true, true, CHK>(a, c1, c2, st); break;
    case 5: launch_bsweep_v<V, R, 64, true, true, CHK>(a, c1, c2, st); break;
    case 6: launch_bsweep_v<V, R, 256, false, true, CHK>(a, c1, c2, st); break;
    default: launch_bsweep_v<V, R, 64, false, true, CHK>(a, c1, c2, st); break;
  }
}

template <typename V, typename R>
static void launch_bsweep_t(const BSweepArgs& a, double omega, int variant, int chk,
                            hipStream_t st) {
  const R c1 = (R)((1.0 + omega - 1) / 1);
  const R c2 = (R)((1.0 + omega - 2) / 2);
  if (chk == 1)
    launch_bsweep_chk<V, R, 1>(a, c1, c2, variant, st);
  else if (chk == 2)
    launch_bsweep_chk<V, R, 2>(a, c1, c2, variant, st);
  else
    launch_bsweep_chk<V, R, 0>(a, c1, c2, variant, st);
}

void launch_batch_sweep(bool fp32, const BatchProblem* probs, const int2* bmap, int nblocks,
                        const int* active, const int* poff, int parity, double omega, int variant,
                        int chk, double* partials, hipStream_t st, int links) {
  const BSweepArgs a{probs, bmap, nblocks, active, poff, parity, partials, links != BLINK_F64};
  if (fp32)
    launch_bsweep_t<float4, float>(a, omega, variant, chk, st);
  else if (links == BLINK_FIXED)
    launch_bsweep_t<int4, double>(a, omega, variant, chk, st);
  else
    launch_bsweep_t<double2, double>(a, omega, variant, chk, st);
}

void launch_batch_finalize(const BatchCheck& c, int nprob, int ndir, double eps,
                           double switch_theta, int maxit, int stall_checks, double stall_gate,
                           hipStream_t st) {
  hipLaunchKernelGGL(k_batch_finalize, dim3(nprob), dim3(BLK), 0, st, c, ndir, eps, switch_theta,
                     maxit, stall_checks, stall_gate);
}

void launch_batch_convert(const BatchProblem* probs32, const BatchProblem* probs64,
                          const int2* bmap, int nblocks, const int* conv, hipStream_t st,
                          bool fixed) {
  if (fixed)
    hipLaunchKernelGGL(k_batch_convert<int4>, dim3(nblocks, 2), dim3(BLK), 0, st, probs32, probs64,
                       bmap, conv);
  else
    hipLaunchKernelGGL(k_batch_convert<double2>, dim3(nblocks, 2), dim3(BLK), 0, st, probs32,
                       probs64, bmap, conv);
}

int batch_convert_block() { return BLK; }

void launch_batch_advance(int* iter, int step, hipStream_t st) {
  hipLaunchKernelGGL(k_batch_advance, dim3(1), dim3(64), 0, st, iter, step);
}

void launch_fill_i32(int* p, int n, int v, hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(k_fill_i32, dim3((n + 255) / 256), dim3(256), 0, st, p, n, v);
}
void launch_fill_f64(double* p, int n, double v, hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(k_fill_f64, dim3((n + 255) / 256), dim3(256), 0, st, p, n, v);
}

int functional_block_size() { return BLK; }

// plain streaming copy (16-byte vectors), the calibration reference for HBM
// byte counters
__global__ __launch_bounds__(256) void k_stream_copy(const float4* __restrict__ src,
                                                     float4* __restrict__ dst, size_t n) {
  for (size_t k = (size_t)blockIdx.x * 256 + threadIdx.x; k < n; k += (size_t)gridDim.x * 256)
    dst[k] = src[k];
}

void launch_stream_copy(const void* src, void* dst, size_t bytes, hipStream_t st) {
  const size_t n = bytes / 16;
  hipLaunchKernelGGL(k_stream_copy, dim3(4096), dim3(256), 0, st, (const float4*)src, (float4*)dst, n);
}

int functional_blocks(int nsh) { return (2 * nsh + BLK - 1) / BLK; }

template <typename V>
static void launch_functional_t(const V* U, const V* ge, const V* go, const Geom& geo, int ndir,
                                int nseg, double* partials, double eps, double* out,
                                hipStream_t st) {
  const int nsh = geo.nh / nseg;
  const int nb = functional_blocks(nsh);
  dim3 grid(nb, nseg);
  if (ndir == 4)
    hipLaunchKernelGGL((k_functional<V, 4>), grid, dim3(BLK), 0, st, U, ge, go, geo, nsh, partials);
  else
    hipLaunchKernelGGL((k_functional<V, 3>), grid, dim3(BLK), 0, st, U, ge, go, geo, nsh, partials);
  const double vseg = 2.0 * nsh;
  hipLaunchKernelGGL(k_finalize, dim3(nseg), dim3(BLK), 0, st, partials, nb, ndir * vseg, vseg,
                     eps, out);
}

void launch_functional_f64(const double2* U, const double2* ge, const double2* go, const Geom& geo,
                           int ndir, int nseg, double* partials, double eps, double* out,
                           hipStream_t st) {
  launch_functional_t<double2>(U, ge, go, geo, ndir, nseg, partials, eps, out, st);
}
void launch_functional_f32(const float4* U, const float4* ge, const float4* go, const Geom& geo,
                           int ndir, int nseg, double* partials, double eps, double* out,
                           hipStream_t st) {
  launch_functional_t<float4>(U, ge, go, geo, ndir, nseg, partials, eps, out, st);
}

void launch_plaquette_f64(const double2* U, const Geom& geo, double* partials, double* out,
                          hipStream_t st) {
  const int nb = (2 * geo.nh + BLK - 1) / BLK;
  hipLaunchKernelGGL(k_plaquette<double2>, dim3(nb), dim3(BLK), 0, st, U, geo, partials);
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(BLK), 0, st, partials, nb, 6.0 * 2.0 * geo.nh,
                     out);
}

void launch_pack_f64(const double* aos, double2* dst, const Geom& geo, int s0, int ns, int nmat,
                     hipStream_t st) {
  FastDiv dL0 = make_fastdiv(geo.L[0]);
  hipLaunchKernelGGL(k_pack<double2>, grid1((long)ns * nmat), dim3(BLK), 0, st, aos, dst, geo, dL0,
                     s0, ns, nmat);
}
void launch_unpack_f64(const double2* src, double* aos, const Geom& geo, int s0, int ns, int nmat,
                       hipStream_t st) {
  FastDiv dL0 = make_fastdiv(geo.L[0]);
  hipLaunchKernelGGL(k_unpack<double2>, grid1((long)ns * nmat), dim3(BLK), 0, st, src, aos, geo,
                     dL0, s0, ns, nmat);
}
void launch_apply_trafo_aos(double* aos, const double2* g, const Geom& geo, int s0, int ns,
                            hipStream_t st) {
  FastDiv dL0 = make_fastdiv(geo.L[0]);
  hipLaunchKernelGGL(k_apply_trafo_aos, grid1((long)ns * 4), dim3(BLK), 0, st, aos, g, geo, dL0, s0,
                     ns);
}
void launch_f64_to_f32(const double2* src, float4* dst, const Geom& geo, int nfield,
                       hipStream_t st) {
  dim3 grid((geo.nh + BLK - 1) / BLK, nfield);
  hipLaunchKernelGGL(k_f64_to_f32, grid, dim3(BLK), 0, st, src, dst, geo.nh, geo.stride, nfield);
}
void launch_pack_links(const double2* src, const PackedArgs& p, const Geom& geo, int nfield,
                       hipStream_t st) {
  dim3 grid((geo.nh + BLK - 1) / BLK, nfield);
  hipLaunchKernelGGL(k_pack_links, grid, dim3(BLK), 0, st, src, (uint4*)p.a, (uint4*)p.b,
                     (uint32_t*)p.c, (uint2*)p.lo, p.a21, p.b21, geo.nh, geo.stride, nfield);
}
void launch_f64_to_fx(const double2* src, int4* dst, const Geom& geo, int nfield, hipStream_t st) {
  dim3 grid((geo.nh + BLK - 1) / BLK, nfield);
  hipLaunchKernelGGL((k_convert<double2, int4, false>), grid, dim3(BLK), 0, st, src, dst, geo.nh,
                     geo.stride);
}
void launch_fx_to_g64(const int4* src, double2* dst, const Geom& geo, hipStream_t st) {
  dim3 grid((geo.nh + BLK - 1) / BLK, 2);
  hipLaunchKernelGGL((k_convert<int4, double2, true>), grid, dim3(BLK), 0, st, src, dst, geo.nh,
                     geo.stride);
}
void launch_g32_to_fx(const float4* src, int4* dst, const Geom& geo, hipStream_t st) {
  dim3 grid((geo.nh + BLK - 1) / BLK, 2);
  hipLaunchKernelGGL((k_convert<float4, int4, true>), grid, dim3(BLK), 0, st, src, dst, geo.nh,
                     geo.stride);
}
void launch_g32_to_g64(const float4* src, double2* dst, const Geom& geo, hipStream_t st) {
  dim3 grid((geo.nh + BLK - 1) / BLK, 2);
  hipLaunchKernelGGL(k_g32_to_g64, grid, dim3(BLK), 0, st, src, dst, geo.nh, geo.stride);
}

void launch_g32_to_fx_range(const float4* se, const float4* so, int4* de, int4* dd,
                            const Geom& geo, int i0, int n, hipStream_t st) {
  dim3 grid((n + BLK - 1) / BLK, 2);
  hipLaunchKernelGGL(k_g32_to_fx_range, grid, dim3(BLK), 0, st, se, so, de, dd, i0, n, geo.stride);
}

void launch_g32_to_g64_range(const float4* se, const float4* so, double2* de, double2* dd,
                             const Geom& geo, int i0, int n, hipStream_t st) {
  dim3 grid((n + BLK - 1) / BLK, 2);
  hipLaunchKernelGGL(k_g32_to_g64_range, grid, dim3(BLK), 0, st, se, so, de, dd, i0, n, geo.stride);
}

}  // namespace gfx
