// gfx_sweep.h -- device code of the overrelaxation sweeps shared by the
// sweep kernels (gfx_kernels.hip) and the persistent form of the batch
// (gfx_persist.hip): block reductions, the site update with its fused
// check, one block of a single-lattice sweep, one block of a batch sweep and
// the per-lattice check of a batch.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "gfx_internal.h"
#include "gfx_lattice.h"
#include "gfx_su3.h"

namespace gfx {

constexpr int BLK = 256;

// ------------------------------------------------------------ reductions
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}

// sums a, b and takes the max of c over the block; result valid in thread 0
__device__ __forceinline__ void block_reduce3(double& a, double& b, double& c) {
  __shared__ double red[3][BLK / 64];
  a = wave_sum(a);
  b = wave_sum(b);
  c = wave_max(c);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) {
    red[0][w] = a;
    red[1][w] = b;
    red[2][w] = c;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nw = blockDim.x >> 6;
    a = red[0][0];
    b = red[1][0];
    c = red[2][0];
    for (int k = 1; k < nw; k++) {
      a += red[0][k];
      b += red[1][k];
      c = fmax(c, red[2][k]);
    }
  }
}

// ------------------------------------------------------------ OR sweep
// Variants (chosen per lattice and precision by the autotuner in gfx_api.hip):
//   BLOCK  workgroup size (256 or 64)
//   NT     non-temporal loads of the links
//   ROLLED the direction loop of the star is kept as a loop: one direction's
//          operands are live at a time (FP64: 162 VGPRs, 3 waves/SIMD).  The
//          other variants: on FP64 and fixed-point links and g the star one
//          term ahead (168 and 164 VGPRs, 3 waves/SIMD); in the FP32 forms
//          the fully unrolled star (every operand in flight, 181 VGPRs, 2
//          waves/SIMD).  In FP32 the rolled variants run the software-
//          pipelined star (star_short_pipelined: two directions' operands in
//          flight, 85-105 VGPRs).
// The FP64 sweep follows the reference's operation order.  The FP32 sweep
// (mixed path, no bitwise target) is bound by its arithmetic as much as by
// HBM, and uses a shorter but equal form: star terms from two rows of each
// SU(3) product, and the overrelaxation series applied to g directly
// (overrelax_apply_short): about 25% fewer complex multiplies.
//
// Fused checks (Landau).  With W_mu(x) = g(x) U_mu(x) g^+(x+mu), the star of a
// site gives g(x) K(x) = sum_mu [W_mu(x) + W_mu(x-mu)^+], so
//   Re tr g(x)K(x) summed over the sites of one parity = sum of Re tr W over
//   all links (every link has one even and one odd end), and
//   Delta(x) = sum_mu [A(W_mu(x-mu)) - A(W_mu(x))] = -A(g(x)K(x))
// with A the traceless hermitian part of overrelax.cu:451-476.  The check
// after iteration i therefore needs no pass of its own:
//   CHK = 1  odd half sweep of iteration i: after the update, g(x)K(x) of the
//            new g(x) -> F and Delta of the odd sites of state i;
//   CHK = 2  even half sweep of iteration i+1: before the update, g(x)K(x)
//            -> Delta of the even sites of state i; the new even field goes
//            to another buffer (gout), so a stop leaves state i intact.
//   CHK = 3  one of the three, given at run time by the argument chk (one
//            body for all three: the persistent form's single call site)
// Each block writes (sum F, sum dAdA, max dAdA) of its sites to
// partials[3 * block].
template <typename R>
__device__ __forceinline__ void site_check(const M3<R>& G, const M3<R>& K, double& F, double& d) {
  M3<R> M = m_mul(G, K);
  F = (double)m_retrace(M);
  vecpot(M);
  R acc = R(0);
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int c = 0; c < 3; c++) acc += M.e[a][c].re * M.e[a][c].re + M.e[a][c].im * M.e[a][c].im;
  d = (double)acc;
}

// LK: the links, as planes of V (const V*) or, for the mixed runs, as
// PackedLinks.  The single-lattice FP64 sweep keeps the reference's
// operation order (on packed links: its star and relaxation, with the series
// applied in the short form); FP32 and, with FSHORT, the batching form's FP64
// sweep use the short form (star_terms_short, overrelax_apply_short).  In
// FP64 on FP64 links the two agree to rounding; where many Gribov copies
// share their links the sweep is bound by its arithmetic, and the short form
// is 10% faster there.  (The short star rebuilds row 2 of each product from
// rows 0 and 1, which packed links, unitary to 1e-11 only, would move by
// more than rounding: the FP64 polish keeps the reference order there.)
template <typename V, typename R, int NDIR, bool ROLLED, bool NT, int CHK, typename LK, bool FSHORT = false>
__device__ __forceinline__ void sweep_site(LK U, const V* __restrict__ gop, const V* gin, V* gout,
                                           const Geom& geo, const Site& s, int i, int parity, R c1,
                                           R c2, double& F, double& d, int chk = CHK) {
  const int S = geo.stride;
  // star K = sum_mu U_mu(x) g^+(x+mu) + U_mu^+(x-mu) g^+(x-mu)
  M3<R> K;
  m_zero(K);
  constexpr bool SHORT = std::is_same<R, float>::value || FSHORT;
  if constexpr (SHORT) {
    // each term is a product of two SU(3) matrices; form its rows 0,1 and
    // rebuild row 2 (U g^+ directly, U^+(x-mu) g^+(x-mu) as (g U)^+)
    if (ROLLED) {
      // FP32: the software-pipelined star (the next direction's operands in
      // flight while a direction's terms are formed); the batching form's
      // FP64 loop loads operand by operand
      if constexpr (std::is_same<R, float>::value) {
        star_short_pipelined<V, R, NT, LK, NDIR>(K, U, gop, geo, s, i, parity);
      } else {
#pragma unroll 1
        for (int mu = 0; mu < NDIR; mu++) star_terms_short<V, R, NT>(K, U, gop, geo, s, i, parity, mu);
      }
    } else {
#pragma unroll
      for (int mu = 0; mu < NDIR; mu++) star_terms_short<V, R, NT>(K, U, gop, geo, s, i, parity, mu);
    }
  } else if (ROLLED) {
    // fixed-point links and g: the star from their integers (raw_m_q)
    constexpr bool FXQ = std::is_same<V, int4>::value && std::is_same<LK, const int4*>::value;
    int jf[NDIR], jb[NDIR];
#pragma unroll
    for (int mu = 0; mu < NDIR; mu++) {
      jf[mu] = nb_fwd(geo, s, i, mu);
      jb[mu] = nb_bwd(geo, s, i, mu);
    }
#pragma unroll 1
    for (int mu = 0; mu < NDIR; mu++) {
      Raw<LK, R> lf, lb;
      Raw<const V*, R> gf, gb;
      raw_load<R, NT>(U, parity * 4 + mu, S, i, lf);
      raw_load<R, false>(gop, S, jf[mu], gf);
      raw_load<R, NT>(U, (1 - parity) * 4 + mu, S, jb[mu], lb);
      raw_load<R, false>(gop, S, jb[mu], gb);
      __builtin_amdgcn_sched_barrier(0);
      if constexpr (FXQ) {
        M3<R> A = raw_m_q<R>(lf);
        M3<R> B = raw_m_dagger_q<R, false>(gf);
        m_addmul(K, A, B);
        A = raw_m_dagger_q<R, true>(lb);
        B = raw_m_dagger_q<R, false>(gb);
        m_addmul(K, A, B);
      } else {
        M3<R> A = raw_m<R>(lf);
        M3<R> B = raw_m_dagger<R>(gf);
        m_addmul(K, A, B);
        A = raw_m_dagger<R>(lb);
        B = raw_m_dagger<R>(gb);
        m_addmul(K, A, B);
      }
    }
    if constexpr (FXQ) fx_star_unscale(K);
  } else if constexpr ((std::is_same<V, int4>::value || std::is_same<V, double2>::value) &&
                       std::is_same<LK, const V*>::value) {
    // fixed-point or FP64 links and g, one star term ahead: the two operands
    // of term t + 1 load while term t is formed (the terms in the rolled
    // loop's order; star_short_pipelined explains the empty asm)
    constexpr bool FXQ = std::is_same<V, int4>::value;
    Raw<LK, R> lr[2];
    Raw<const V*, R> gr[2];
    auto load_term = [&](int t, int k, int z) {
      const int mu = t >> 1;
      if ((t & 1) == 0) {
        raw_load<R, NT>(U, parity * 4 + mu, S, i + z, lr[k]);
        raw_load<R, false>(gop, S, nb_fwd(geo, s, i, mu) + z, gr[k]);
      } else {
        const int jb = nb_bwd(geo, s, i, mu) + z;
        raw_load<R, NT>(U, (1 - parity) * 4 + mu, S, jb, lr[k]);
        raw_load<R, false>(gop, S, jb, gr[k]);
      }
    };
    load_term(0, 0, 0);
#pragma unroll
    for (int t = 0; t < 2 * NDIR; t++) {
      const int k = t & 1;
      if (t + 1 < 2 * NDIR) {
        int ii = i;
        asm volatile(""
                     : "+v"(ii), "+v"(K.e[0][0].re), "+v"(K.e[0][0].im), "+v"(K.e[0][1].re),
                       "+v"(K.e[0][1].im), "+v"(K.e[0][2].re), "+v"(K.e[0][2].im),
                       "+v"(K.e[1][0].re), "+v"(K.e[1][0].im), "+v"(K.e[1][1].re),
                       "+v"(K.e[1][1].im), "+v"(K.e[1][2].re), "+v"(K.e[1][2].im),
                       "+v"(K.e[2][0].re), "+v"(K.e[2][0].im), "+v"(K.e[2][1].re),
                       "+v"(K.e[2][1].im), "+v"(K.e[2][2].re), "+v"(K.e[2][2].im));
        load_term(t + 1, 1 - k, ii - i);
      }
      if constexpr (FXQ) {
        const M3<R> A = k ? raw_m_dagger_q<R, true>(lr[k]) : raw_m_q<R>(lr[k]);
        const M3<R> B = raw_m_dagger_q<R, false>(gr[k]);
        m_addmul(K, A, B);
      } else {
        const M3<R> A = k ? raw_m_dagger<R>(lr[k]) : raw_m<R>(lr[k]);
        const M3<R> B = raw_m_dagger<R>(gr[k]);
        m_addmul(K, A, B);
      }
    }
    if constexpr (FXQ) fx_star_unscale(K);
  } else {
#pragma unroll
    for (int mu = 0; mu < NDIR; mu++) {
      const int jf = nb_fwd(geo, s, i, mu);
      const int jb = nb_bwd(geo, s, i, mu);
      M3<R> A = link_m<R, NT>(U, parity * 4 + mu, S, i);
      M3<R> B = load_m_dagger<V, R>(gop, S, jf);
      m_addmul(K, A, B);
      A = link_m_dagger<R, NT>(U, (1 - parity) * 4 + mu, S, jb);
      B = load_m_dagger<V, R>(gop, S, jb);
      m_addmul(K, A, B);
    }
  }
  M3<R> G = load_m<V, R>(gin, S, i);
  if (CHK == 2 || (CHK == 3 && chk == 2)) site_check(G, K, F, d);
  M3<R> h = G;
  cm_relax(h, K);
  // the polish on packed or fixed-point links applies the series in the
  // short form too: its star and relaxation keep the reference's order,
  // which is what pins the packing to the oracle; the series differs from
  // it by rounding only.  The short form takes h as the relaxation leaves
  // it: h is G rotated by three normalised SU(2) elements, as unitary as G,
  // and the series' result is normalised anyway (the reference normalises
  // h first, overrelax.cu:413)
  if constexpr (SHORT || std::is_same<LK, PackedLinks>::value || std::is_same<V, int4>::value) {
    overrelax_apply_short(G, h, c1, c2);
  } else {
    m_normalize(h);
    overrelax_apply(G, h, c1, c2);
    m_normalize(G);
  }
  if (CHK == 1 || (CHK == 3 && chk == 1)) site_check(G, K, F, d);
  Store<V>::store(gout, S, i, G);
}

template <int CHK>
__device__ __forceinline__ void store_check_partials(double F, double d, double* partials, int slot,
                                                     int chk = CHK) {
  double mx = d;
  block_reduce3(F, d, mx);
  if (threadIdx.x == 0) {
    double* o = partials + 3 * (size_t)slot;
    o[0] = (CHK == 1 || (CHK == 3 && chk == 1)) ? F : 0.0;
    o[1] = d;
    o[2] = mx;
  }
}

// ge / go: even and odd fields of g.  The new own field goes to galt when
// given (always with CHK = 2), otherwise the sweep is in place.
// bmap[launch position] = block of the lattice (workgroup order)
template <typename V, typename R, int NDIR, int BLOCK, bool ROLLED, bool NT, int CHK, typename LK>
__device__ __forceinline__ void or_sweep_block(LK U, V* ge, V* go, V* galt, const Geom& geo,
                                               int parity, R c1, R c2,
                                               const unsigned char* __restrict__ active,
                                               const int* __restrict__ bmap,
                                               double* __restrict__ partials) {
  const int b = bmap[blockIdx.x];
  const int i = b * BLOCK + threadIdx.x;
  V* gme = parity ? go : ge;
  const V* gop = parity ? ge : go;
  V* gout = (galt != nullptr) ? galt : gme;
  double F = 0.0, d = 0.0;
  if (CHK == 0) {
    if (i >= geo.nh) return;
    const Site s = site_coords(geo, i, parity);
    if (NDIR == 3 && active != nullptr && !active[s.x[3]]) return;
    sweep_site<V, R, NDIR, ROLLED, NT, 0>(U, gop, gme, gout, geo, s, i, parity, c1, c2, F, d);
  } else {
    if (i < geo.nh) {
      const Site s = site_coords(geo, i, parity);
      // Coulomb: a block lies in one time slice (slices hold a whole number
      // of blocks when checks are fused), so the whole block leaves together
      if (NDIR == 3 && !active[s.x[3]]) return;
      sweep_site<V, R, NDIR, ROLLED, NT, CHK>(U, gop, gme, gout, geo, s, i, parity, c1, c2, F, d);
    }
    store_check_partials<CHK>(F, d, partials, b);
  }
}

// ------------------------------------------------------------ batched sweep
// Many lattices (Gribov copies of one configuration sharing its links, or
// configurations of different sizes) in one launch.  bmap[block] = (problem,
// block inside the problem); a problem whose active flag is 0 has converged
// and its blocks return at once.  Check partials of block (q, k) go to
// partials[3 * (poff[q] + k)].  FP32 and FP64 on FP64 or packed links use
// the short form of the products (see sweep_site); on fixed-point links and
// g (V = int4) the star and relaxation keep the reference's order, as in the
// single-lattice polish, which the fixed-point stop bound assumes.  PK: the
// links are read from the packed planes P.pk (FP32: the 21-bit planes,
// FP64: high + low parts).
// batch_sweep_item: one block of the map, m = (problem, block inside the
// problem).  Also the work item of the persistent form (k_batch_persistent,
// CHK = 3 with the check form in chk), whose per-lattice check changes
// probs, active and the partials between the items it runs: no restrict
// qualifiers here.
template <typename V, typename R, int NDIR, int BLOCK, bool ROLLED, bool NT, int CHK, bool PK>
__device__ __forceinline__ void batch_sweep_item(const BatchProblem* probs, int2 m,
                                                 const int* active, const int* poff, int parity,
                                                 R c1, R c2, double* partials, int chk = CHK) {
  if (!active[m.x]) return;
  const BatchProblem& P = probs[m.x];
  const Geom& geo = P.geo;
  const V* U = static_cast<const V*>(P.U);
  V* gme = static_cast<V*>(P.g[parity]);
  const V* gop = static_cast<const V*>(P.g[1 - parity]);
  V* gout = (CHK == 2 || (CHK == 3 && chk == 2)) ? static_cast<V*>(P.galt) : gme;
  const int i = m.y * BLOCK + threadIdx.x;
  double F = 0.0, d = 0.0;
  constexpr bool FSHORT = !std::is_same<V, int4>::value;
  auto site = [&](const Site& s) {
    if constexpr (PK)
      sweep_site<V, R, NDIR, ROLLED, NT, CHK, PackedLinks, true>(P.pk, gop, gme, gout, geo, s, i,
                                                                 parity, c1, c2, F, d, chk);
    else
      sweep_site<V, R, NDIR, ROLLED, NT, CHK, const V*, FSHORT>(U, gop, gme, gout, geo, s, i,
                                                                parity, c1, c2, F, d, chk);
  };
  if (CHK == 0 || (CHK == 3 && chk == 0)) {
    if (i >= geo.nh) return;
    site(site_coords(geo, i, parity));
  } else {
    if (i < geo.nh) site(site_coords(geo, i, parity));
    store_check_partials<CHK>(F, d, partials, poff[m.x] + m.y, chk);
  }
}

// The check of one lattice q of a batch at the end of a check period (the
// body of k_batch_finalize, one workgroup per lattice): its block partials
// summed, F, dAdA and max dAdA recorded, and the lattice stopped, moved from
// FP32 to FP64, or its fields swapped to go on.  `changed` (when given) is
// set when a lattice stops or leaves FP32.
__device__ __forceinline__ void batch_finalize_lattice(int q, const BatchCheck& c, int ndir,
                                                       double eps, double switch_theta, int maxit,
                                                       int stall_checks, double stall_gate,
                                                       int* changed) {
  const bool f64 = c.act[0][q] != 0;
  if (!f64 && !c.act[1][q]) return;
  // a lattice whose blocks are not in its precision's workgroup map (it
  // left FP32 inside a graph captured before) was not swept: it waits
  if (!(f64 ? c.inmap[0][q] : c.inmap[1][q])) return;
  const double* pod = f64 ? c.pod[0] : c.pod[1];
  const double* pev = f64 ? c.pev[0] : c.pev[1];
  const int* poff = f64 ? c.poff[0] : c.poff[1];
  double F = 0.0, d = 0.0, mx = 0.0;
#pragma unroll 8
  for (int b = poff[q] + threadIdx.x; b < poff[q + 1]; b += blockDim.x) {
    F += pod[3 * (size_t)b];
    d += pod[3 * (size_t)b + 1] + pev[3 * (size_t)b + 1];
    mx = fmax(mx, fmax(pod[3 * (size_t)b + 2], pev[3 * (size_t)b + 2]));
  }
  block_reduce3(F, d, mx);
  if (threadIdx.x != 0) return;
  BatchProblem* probs64 = c.probs[0];
  BatchProblem* probs32 = c.probs[1];
  const double V = 2.0 * probs64[q].geo.nh;
  const int it = __hip_atomic_load(c.iter, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const bool last = it + 1 >= maxit;
  double* results = c.results;
  results[4 * q + 0] = F / (ndir * V);
  results[4 * q + 1] = d / V;
  results[4 * q + 2] = mx;
  // on packed or fixed-point fields an FP64 lattice stops when the
  // rounding's worst effect on Delta (packed links: with the short star of
  // the batch sweeps; fixed-point links and g: reference-order star, g
  // reunitarised at the end) keeps max dAdA below eps on the exact links
  // and the final FP64 g as well
  if (f64) {
    const int links = c.links;
    const double bound = links == 1 ? PACK_DDELTA64_SHORT : links == 2 ? FX_DDELTA : 0.0;
    const double r = sqrt(mx) + bound;
    if (links ? r * r < eps : mx < eps) {
      results[4 * q + 3] = (double)it;
      c.act[0][q] = 0;
      if (changed) __hip_atomic_store(changed, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else if (!last) {
      void* t = probs64[q].g[0];
      probs64[q].g[0] = probs64[q].galt;
      probs64[q].galt = t;
    }
    return;
  }
  // an FP32 lattice never stops: below eps (or the switch, or at its stall)
  // it moves to FP64, whose check decides its stop
  bool leave = mx < eps || mx < switch_theta;
  if (!leave && stall_checks > 0) {
    if (mx < 0.5 * c.best[q]) {
      c.best[q] = mx;
      c.stall[q] = 0;
    } else if (c.best[q] < stall_gate && ++c.stall[q] >= stall_checks) {
      leave = true;
    }
  }
  if (leave) {
    c.act[1][q] = 0;
    c.sw[q] = it;
    if (last) {
      c.conv[q] = 2;
    } else {
      c.conv[q] = 1;
      c.act[0][q] = 1;
    }
    if (changed) __hip_atomic_store(changed, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  } else if (!last) {
    void* t = probs32[q].g[0];
    probs32[q].g[0] = probs32[q].galt;
    probs32[q].galt = t;
  }
}

}  // namespace gfx
