// gfx_heatbath.hip -- stochastic updates on the checkerboard:
//   k_sa_sweep      simulated annealing of g: Cabibbo-Marinari heatbath of
//                   exp(Re tr(g K)/T) (simulated_annealing.cu:5-257)
//   k_micro_sweep   microcanonical (F-preserving) reflection of g in each
//                   SU(2) subgroup (heatbath_thermalization.cu:318-401,
//                   applied to g)
//   k_link_heatbath SU(3) link heatbath of the Wilson action with the
//                   staple of heatbath_thermalization.cu:235-308
//   k_batch_sa_sweep / k_batch_micro_sweep   the same updates of g over
//                   every lattice of a batch in one launch
//   k_haar_trafo    Haar-random g from the twisters' normals
//
// Two SU(2) samplers:
//   SAMPLER_REF  the reference's rule (simulated_annealing.cu:5-73): one
//                trial with a0 = 1 - (g^2 - ln(1-u0))/(|w|^2 beta), accept if
//                2 u1^2 <= 2 - y, otherwise alpha = w (unnormalised); the same
//                5 numbers serve all three subgroups.  Random layout
//                gauss[i], unif[4i+k] with i the parity index.
//   SAMPLER_KP   Kennedy-Pendleton (Phys. Lett. B156 (1985) 393): per trial
//                x = (g^2/2 - ln u)/(2 a), accept if v^2 <= 1 - x, a0 = 1 - 2x,
//                a = beta*|w|; up to NT trials per subgroup, the subgroup is
//                left unchanged if all fail (a mixture of the identity and
//                the exact conditional, so the Boltzmann weight stays
//                invariant).  Layout gauss[NG*i + sub*NT + t],
//                unif[NU*i + sub*(2NT+2) + ...].
#include <hip/hip_runtime.h>

#include "gfx_lattice.h"
#include "gfx_su3.h"

namespace gfx {

constexpr int HB_BLK = 256;
constexpr int KP_NT = 4;
constexpr int KP_NG = 3 * KP_NT;
constexpr int KP_NU = 3 * (2 * KP_NT + 2);

// quaternion product in the (a.x, a.y, b.x, b.y) component order of the
// reference (dev_su3.cu:1385-1399); equals the product of the 2x2 matrices
// [[q0 + i q3, q2 + i q1], [-q2 + i q1, q0 - i q3]]
__device__ __forceinline__ void su2_mul(double r[4], const double a[4], const double b[4]) {
  r[0] = a[0] * b[0] - a[1] * b[1] - a[2] * b[2] - a[3] * b[3];
  r[1] = a[0] * b[1] + a[1] * b[0] - a[2] * b[3] + a[3] * b[2];
  r[2] = a[0] * b[2] + a[2] * b[0] - a[3] * b[1] + a[1] * b[3];
  r[3] = a[0] * b[3] + a[3] * b[0] - a[1] * b[2] + a[2] * b[1];
}

// simulated_annealing.cu:5-73 restated
__device__ __forceinline__ void sample_ref(double alpha[4], const double w[4], double beta,
                                           const float* __restrict__ gauss,
                                           const float* __restrict__ unif, int ind) {
  const double k = w[0] * w[0] + w[1] * w[1] + w[2] * w[2] + w[3] * w[3];
  const double rdet = ::rsqrt(k);
  const double v[4] = {w[0] * rdet, w[1] * rdet, w[2] * rdet, w[3] * rdet};
  const double kb = k * beta;
  float a = fabsf(gauss[ind]);
  float b = -logf(1.0f - unif[4 * ind]);
  float y = a * a + b;
  y = (float)((double)y / kb);
  const float help = unif[4 * ind + 1];
  double q[4];
  int ret;
  if ((2.0f * (help * help)) <= (2.0f - y)) {
    q[0] = (double)(1.0f - y);
    ret = 1;
  } else {
    q[0] = 0.0;
    ret = 0;
  }
  const double norm = sqrt(1.0 - q[0] * q[0]);
  const double ct = 2.0 * (double)unif[4 * ind + 2] - 1.0;
  const double st = sqrt(1.0 - ct * ct);
  const double phi = 6.2831853071795862 * (double)unif[4 * ind + 3];
  double sp, cp;
  sincos(phi, &sp, &cp);
  q[1] = norm * st * cp;
  q[2] = norm * st * sp;
  q[3] = norm * ct;
  if (ret == 1) {
    su2_mul(alpha, q, v);
  } else {
    alpha[0] = w[0];
    alpha[1] = w[1];
    alpha[2] = w[2];
    alpha[3] = w[3];
  }
}

// Kennedy-Pendleton; returns false if every trial was rejected
__device__ __forceinline__ bool sample_kp(double alpha[4], const double w[4], double beta,
                                          const float* __restrict__ g, const float* __restrict__ u) {
  const double k = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2] + w[3] * w[3]);
  if (!(k > 1e-300)) return false;
  const double a = beta * k;
  double a0 = 0.0;
  bool ok = false;
#pragma unroll
  for (int t = 0; t < KP_NT; t++) {
    if (!ok) {
      const double gg = (double)g[t];
      const double x = (0.5 * gg * gg - log((double)u[2 * t])) / (2.0 * a);
      const double acc = (double)u[2 * t + 1];
      if (x <= 1.0 && acc * acc <= 1.0 - x) {
        a0 = 1.0 - 2.0 * x;
        ok = true;
      }
    }
  }
  if (!ok) return false;
  const double norm = sqrt(fmax(0.0, 1.0 - a0 * a0));
  const double ct = 2.0 * (double)u[2 * KP_NT] - 1.0;
  const double st = sqrt(fmax(0.0, 1.0 - ct * ct));
  const double phi = 6.283185307179586 * (double)u[2 * KP_NT + 1];
  double sp, cp;
  sincos(phi, &sp, &cp);
  const double q[4] = {a0, norm * st * cp, norm * st * sp, norm * ct};
  const double v[4] = {w[0] / k, w[1] / k, w[2] / k, w[3] / k};
  su2_mul(alpha, q, v);
  return true;
}

// one heatbath update of h in each SU(2) subgroup; the subgroup vector w
// comes from h K in the storage arithmetic R, the sampling runs in FP64
// (F: products through fused multiply-adds, see gfx_su3.h)
template <int SAMPLER, typename R, bool F = true>
__device__ __forceinline__ void cm_heatbath(M3<R>& h, const M3<R>& K, double beta,
                                            const float* __restrict__ gauss,
                                            const float* __restrict__ unif, int i) {
  int sub = 0;
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = a + 1; b < 3; b++) {
      R wr[4];
      double w[4], al[4];
      su2_subvec<R, F>(h, K, a, b, wr[0], wr[1], wr[2], wr[3]);
#pragma unroll
      for (int c = 0; c < 4; c++) w[c] = (double)wr[c];
      if (SAMPLER == 0) {
        sample_ref(al, w, beta, gauss, unif, i);
        su2_left<R, F>(h, a, b, (R)al[0], (R)al[1], (R)al[2], (R)al[3]);
      } else {
        const float* g = gauss + (size_t)KP_NG * i + sub * KP_NT;
        const float* u = unif + (size_t)KP_NU * i + sub * (2 * KP_NT + 2);
        if (sample_kp(al, w, beta, g, u))
          su2_left<R, F>(h, a, b, (R)al[0], (R)al[1], (R)al[2], (R)al[3]);
      }
      sub++;
    }
}

// star K of g at site i (same as the overrelaxation sweep).  NT: the links
// are loaded non-temporally (a single lattice reads each once per half
// sweep, and L2 keeps the neighbour g instead); Gribov copies sharing their
// links load them through the caches
template <int NDIR, bool NT, typename V = double2, typename R = double>
__device__ __forceinline__ M3<R> star_of(const V* __restrict__ U, const V* __restrict__ gop,
                                         const Geom& geo, const Site& s, int i, int parity) {
  M3<R> K;
  m_zero(K);
  // direction loop kept rolled (one direction's operands live at a time:
  // unrolled, these kernels spilled to scratch at 1 wave/SIMD), the four
  // operands of a direction loaded before any is decoded, products in the
  // short form (the stochastic updates have no bitwise target)
  // (the software-pipelined star of the FP32 OR sweeps takes the batched
  // anneal's FP32 kernels from 88 to 107-126 VGPRs, 4 instead of 5 waves per
  // SIMD: with the links shared by the copies and served from L2 they are
  // bound by their arithmetic, and the 4000-step anneal took 23.8 instead
  // of 23.1-23.4 s)
#pragma unroll 1
  for (int mu = 0; mu < NDIR; mu++)
    star_terms_short<V, R, NT, const V*, true>(K, U, gop, geo, s, i, parity, mu);
  return K;
}

// bmap: launch position -> block of HB_BLK sites (a sweep workgroup order
// dealt to the XCDs); the sites of one parity are updated from the other
// parity only and draw the numbers of their own index, so the order leaves
// the result unchanged
template <int SAMPLER, int NDIR>
__global__ __launch_bounds__(HB_BLK) void k_sa_sweep(const double2* __restrict__ U,
                                                     double2* __restrict__ g, Geom geo, int parity,
                                                     double beta, const float* __restrict__ gauss,
                                                     const float* __restrict__ unif,
                                                     const int* __restrict__ bmap) {
  const int i = bmap[blockIdx.x] * HB_BLK + threadIdx.x;
  if (i >= geo.nh) return;
  const Site s = site_coords(geo, i, parity);
  const int S = geo.stride;
  const M3<double> K = star_of<NDIR, true>(U, field_ptr(static_cast<const double2*>(g), 1 - parity, S),
                                     geo, s, i, parity);
  double2* gme = field_ptr(g, parity, S);
  M3<double> h = load_m<double2, double>(gme, S, i);
  cm_heatbath<SAMPLER>(h, K, beta, gauss, unif, i);
  m_normalize(h);
  Store<double2>::store(gme, S, i, h);
}

// microcanonical reflection in each SU(2) subgroup:
// heatbath_thermalization.cu:318-338, alpha = 2 w0 w/|w|^2 - e0
template <typename R>
__device__ __forceinline__ void micro_reflect(M3<R>& h, const M3<R>& K) {
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = a + 1; b < 3; b++) {
      R w[4];
      su2_subvec(h, K, a, b, w[0], w[1], w[2], w[3]);
      const R ww = w[0] * w[0] + w[1] * w[1] + w[2] * w[2] + w[3] * w[3];
      if (!(fabs(ww) < R(1.0e-9))) {
        const R f = R(2) * w[0] / ww;
        su2_left(h, a, b, f * w[0] - R(1), f * w[1], f * w[2], f * w[3]);
      }
    }
}

template <int NDIR>
__global__ __launch_bounds__(HB_BLK) void k_micro_sweep(const double2* __restrict__ U,
                                                        double2* __restrict__ g, Geom geo,
                                                        int parity, const int* __restrict__ bmap) {
  const int i = bmap[blockIdx.x] * HB_BLK + threadIdx.x;
  if (i >= geo.nh) return;
  const Site s = site_coords(geo, i, parity);
  const int S = geo.stride;
  const M3<double> K = star_of<NDIR, true>(U, field_ptr(static_cast<const double2*>(g), 1 - parity, S),
                                     geo, s, i, parity);
  double2* gme = field_ptr(g, parity, S);
  M3<double> h = load_m<double2, double>(gme, S, i);
  micro_reflect(h, K);
  m_normalize(h);
  Store<double2>::store(gme, S, i, h);
}

// ---- batched annealing of many lattices (Gribov copies, batches):
// bmap[block] = (lattice, block of HB_BLK sites of the lattice); the random
// numbers of lattice q start at site roff[q] of the fields; the inverse
// temperature of the current step is betas[*step] (step counter on the
// device, so a captured graph of several steps replays unchanged).
template <int SAMPLER, typename V, typename R>
__global__ __launch_bounds__(HB_BLK) void k_batch_sa_sweep(const BatchProblem* __restrict__ probs,
                                                           const int2* __restrict__ bmap,
                                                           const int* __restrict__ roff, int parity,
                                                           const double* __restrict__ betas,
                                                           const int* __restrict__ step,
                                                           const float* __restrict__ gauss,
                                                           const float* __restrict__ unif) {
  const int2 m = bmap[blockIdx.x];
  const BatchProblem& P = probs[m.x];
  const Geom& geo = P.geo;
  const int i = m.y * HB_BLK + threadIdx.x;
  if (i >= geo.nh) return;
  const Site s = site_coords(geo, i, parity);
  V* gme = static_cast<V*>(P.g[parity]);
  const M3<R> K = star_of<4, false, V, R>(static_cast<const V*>(P.U), static_cast<const V*>(P.g[1 - parity]),
                                   geo, s, i, parity);
  M3<R> h = load_m<V, R>(gme, geo.stride, i);
  cm_heatbath<SAMPLER>(h, K, betas[*step], gauss, unif, roff[m.x] + i);
  m_normalize(h);
  Store<V>::store(gme, geo.stride, i, h);
}

template <typename V, typename R>
__global__ __launch_bounds__(HB_BLK) void k_batch_micro_sweep(const BatchProblem* __restrict__ probs,
                                                              const int2* __restrict__ bmap,
                                                              int parity) {
  const int2 m = bmap[blockIdx.x];
  const BatchProblem& P = probs[m.x];
  const Geom& geo = P.geo;
  const int i = m.y * HB_BLK + threadIdx.x;
  if (i >= geo.nh) return;
  const Site s = site_coords(geo, i, parity);
  V* gme = static_cast<V*>(P.g[parity]);
  const M3<R> K = star_of<4, false, V, R>(static_cast<const V*>(P.U), static_cast<const V*>(P.g[1 - parity]),
                                   geo, s, i, parity);
  M3<R> h = load_m<V, R>(gme, geo.stride, i);
  micro_reflect(h, K);
  m_normalize(h);
  Store<V>::store(gme, geo.stride, i, h);
}

// staple of U_nu(x): sum_{mu != nu} U_mu(x+nu) U_nu^+(x+mu) U_mu^+(x)
//                                 + U_mu^+(x+nu-mu) U_nu^+(x-mu) U_mu(x-mu)
// (the products in their earlier, unfused form: F = false, so that a
// thermalized configuration stays the same for the same twister seeds)
// bmap: launch position -> block (a sweep workgroup order of HB_BLK-site
// blocks, dealt to the XCDs); the update of a link depends only on links of
// the other parity or of other directions, so the order leaves the result
// unchanged and only decides how much of the staples' links L2 serves
__global__ __launch_bounds__(HB_BLK) void k_link_heatbath(double2* __restrict__ U, Geom geo,
                                                          int parity, int nu, double beta,
                                                          const float* __restrict__ gauss,
                                                          const float* __restrict__ unif,
                                                          const int* __restrict__ bmap) {
  const int i = bmap[blockIdx.x] * HB_BLK + threadIdx.x;
  if (i >= geo.nh) return;
  const int S = geo.stride;
  const int p = parity, q = 1 - parity;
  const Site s = site_coords(geo, i, p);
  const int xpnu = nb_fwd(geo, s, i, nu);  // parity q
  const Site snu = site_coords(geo, xpnu, q);
  M3<double> st;
  m_zero(st);
#pragma unroll
  for (int mu = 0; mu < 4; mu++) {
    if (mu == nu) continue;
    const int xpmu = nb_fwd(geo, s, i, mu);        // parity q
    const int xmmu = nb_bwd(geo, s, i, mu);        // parity q
    const int xpnumu = nb_bwd(geo, snu, xpnu, mu); // x+nu-mu, parity p
    // forward staple
    M3<double> a = load_m<double2, double, false, false>(field_ptr(U, q * 4 + mu, S), S, xpnu);
    M3<double> b = load_m_dagger<double2, double, false, false>(field_ptr(U, q * 4 + nu, S), S, xpmu);
    M3<double> c = m_mul<double, false>(a, b);
    M3<double> d = load_m_dagger<double2, double, false, false>(field_ptr(U, p * 4 + mu, S), S, i);
    m_addmul<double, false>(st, c, d);
    // backward staple
    a = load_m_dagger<double2, double, false, false>(field_ptr(U, p * 4 + mu, S), S, xpnumu);
    b = load_m_dagger<double2, double, false, false>(field_ptr(U, q * 4 + nu, S), S, xmmu);
    c = m_mul<double, false>(a, b);
    d = load_m<double2, double, false, false>(field_ptr(U, q * 4 + mu, S), S, xmmu);
    m_addmul<double, false>(st, c, d);
  }
  double2* Un = field_ptr(U, p * 4 + nu, S);
  M3<double> h = load_m<double2, double, false, false>(Un, S, i);
  cm_heatbath<1, double, false>(h, st, beta, gauss, unif, i);
  m_normalize<double, false>(h);
  Store<double2>::store(Un, S, i, h);
}

// Haar-random SU(3) g(x): rows 0,1 from 12 standard normals per site (two
// complex Gaussian vectors), Gram-Schmidt, row 2 = conj(row0 x row1).  The
// first two rows of a Haar unitary matrix are two orthonormalised complex
// Gaussian vectors, and fixing row 2 this way puts det = 1.
__global__ __launch_bounds__(HB_BLK) void k_haar_trafo(double2* __restrict__ g, Geom geo,
                                                       FastDiv dL0, const float* __restrict__ gauss) {
  const int s = blockIdx.x * HB_BLK + threadIdx.x;
  if (s >= 2 * geo.nh) return;
  uint32_t q = fdiv((uint32_t)s, dL0);
  const int x0 = s - (int)q * geo.L[0];
  uint32_t q2 = fdiv(q, geo.dL1);
  const int x1 = (int)(q - q2 * (uint32_t)geo.L[1]);
  uint32_t q3 = fdiv(q2, geo.dL2);
  const int x2 = (int)(q2 - q3 * (uint32_t)geo.L[2]);
  const int p = (x0 + x1 + x2 + (int)q3) & 1;
  const float* z = gauss + 12 * (size_t)s;
  M3<double> m;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    m.e[0][c] = cx<double>((double)z[2 * c], (double)z[2 * c + 1]);
    m.e[1][c] = cx<double>((double)z[6 + 2 * c], (double)z[6 + 2 * c + 1]);
  }
  m_normalize<double, false>(m);  // earlier, unfused form: the same starts as before
  Store<double2>::store(field_ptr(g, p, geo.stride), geo.stride, s >> 1, m);
}

void launch_haar_trafo(double2* g, const Geom& geo, const float* gauss, hipStream_t st) {
  hipLaunchKernelGGL(k_haar_trafo, dim3((2 * geo.nh + HB_BLK - 1) / HB_BLK), dim3(HB_BLK), 0, st,
                     g, geo, make_fastdiv(geo.L[0]), gauss);
}

void launch_sa_sweep(const double2* U, double2* g, const Geom& geo, int parity, int ndir,
                     double beta, int sampler, const float* gauss, const float* unif,
                     const int* bmap, hipStream_t st) {
  dim3 grid((geo.nh + HB_BLK - 1) / HB_BLK);
  if (sampler == 0) {
    if (ndir == 4)
      hipLaunchKernelGGL((k_sa_sweep<0, 4>), grid, dim3(HB_BLK), 0, st, U, g, geo, parity, beta, gauss, unif, bmap);
    else
      hipLaunchKernelGGL((k_sa_sweep<0, 3>), grid, dim3(HB_BLK), 0, st, U, g, geo, parity, beta, gauss, unif, bmap);
  } else {
    if (ndir == 4)
      hipLaunchKernelGGL((k_sa_sweep<1, 4>), grid, dim3(HB_BLK), 0, st, U, g, geo, parity, beta, gauss, unif, bmap);
    else
      hipLaunchKernelGGL((k_sa_sweep<1, 3>), grid, dim3(HB_BLK), 0, st, U, g, geo, parity, beta, gauss, unif, bmap);
  }
}

void launch_micro_sweep(const double2* U, double2* g, const Geom& geo, int parity, int ndir,
                        const int* bmap, hipStream_t st) {
  dim3 grid((geo.nh + HB_BLK - 1) / HB_BLK);
  if (ndir == 4)
    hipLaunchKernelGGL((k_micro_sweep<4>), grid, dim3(HB_BLK), 0, st, U, g, geo, parity, bmap);
  else
    hipLaunchKernelGGL((k_micro_sweep<3>), grid, dim3(HB_BLK), 0, st, U, g, geo, parity, bmap);
}

void launch_link_heatbath(double2* U, const Geom& geo, int parity, int nu, double beta,
                          const float* gauss, const float* unif, const int* bmap, hipStream_t st) {
  dim3 grid((geo.nh + HB_BLK - 1) / HB_BLK);
  hipLaunchKernelGGL(k_link_heatbath, grid, dim3(HB_BLK), 0, st, U, geo, parity, nu, beta, gauss,
                     unif, bmap);
}

// the SU(2) samplers alone, for distribution tests: sample i draws alpha
// for w = (wnorm, 0, 0, 0) from the random fields exactly as subgroup 0 of
// site i in k_sa_sweep; ALL: out[4i..4i+3] = alpha (quaternion components
// in the reference's order), else out[i] = alpha_0 (NaN when
// Kennedy-Pendleton rejected every trial and the subgroup would stay
// unchanged)
template <int SAMPLER, bool ALL>
__global__ __launch_bounds__(HB_BLK) void k_su2_sample(double beta, double wnorm, int n,
                                                       const float* __restrict__ gauss,
                                                       const float* __restrict__ unif,
                                                       double* __restrict__ out) {
  const int i = blockIdx.x * HB_BLK + threadIdx.x;
  if (i >= n) return;
  const double w[4] = {wnorm, 0.0, 0.0, 0.0};
  double al[4];
  bool ok = true;
  if (SAMPLER == 0)
    sample_ref(al, w, beta, gauss, unif, i);
  else
    ok = sample_kp(al, w, beta, gauss + (size_t)KP_NG * i, unif + (size_t)KP_NU * i);
  if (ALL) {
#pragma unroll
    for (int k = 0; k < 4; k++) out[4 * (size_t)i + k] = ok ? al[k] : __builtin_nan("");
  } else {
    out[i] = ok ? al[0] : __builtin_nan("");
  }
}

void launch_su2_sample(int sampler, double beta, double wnorm, int n, const float* gauss,
                       const float* unif, double* out, hipStream_t st, bool all) {
  const dim3 grid((n + HB_BLK - 1) / HB_BLK);
  if (sampler == 0) {
    if (all)
      hipLaunchKernelGGL((k_su2_sample<0, true>), grid, dim3(HB_BLK), 0, st, beta, wnorm, n, gauss, unif, out);
    else
      hipLaunchKernelGGL((k_su2_sample<0, false>), grid, dim3(HB_BLK), 0, st, beta, wnorm, n, gauss, unif, out);
  } else {
    if (all)
      hipLaunchKernelGGL((k_su2_sample<1, true>), grid, dim3(HB_BLK), 0, st, beta, wnorm, n, gauss, unif, out);
    else
      hipLaunchKernelGGL((k_su2_sample<1, false>), grid, dim3(HB_BLK), 0, st, beta, wnorm, n, gauss, unif, out);
  }
}

template <typename V, typename R>
static void launch_bsa(const BatchProblem* probs, const int2* bmap, int nblocks, const int* roff,
                       int parity, const double* betas, const int* step, int sampler,
                       const float* gauss, const float* unif, hipStream_t st) {
  if (sampler == 0)
    hipLaunchKernelGGL((k_batch_sa_sweep<0, V, R>), dim3(nblocks), dim3(HB_BLK), 0, st, probs, bmap,
                       roff, parity, betas, step, gauss, unif);
  else
    hipLaunchKernelGGL((k_batch_sa_sweep<1, V, R>), dim3(nblocks), dim3(HB_BLK), 0, st, probs, bmap,
                       roff, parity, betas, step, gauss, unif);
}

// fp32: the lattices' FP32 problems (FP32 links and g); else FP64
void launch_batch_sa_sweep(const BatchProblem* probs, const int2* bmap, int nblocks,
                           const int* roff, int parity, const double* betas, const int* step,
                           int sampler, const float* gauss, const float* unif, hipStream_t st,
                           bool fp32) {
  if (fp32)
    launch_bsa<float4, float>(probs, bmap, nblocks, roff, parity, betas, step, sampler, gauss, unif, st);
  else
    launch_bsa<double2, double>(probs, bmap, nblocks, roff, parity, betas, step, sampler, gauss, unif,
                                st);
}

void launch_batch_micro_sweep(const BatchProblem* probs, const int2* bmap, int nblocks, int parity,
                              hipStream_t st, bool fp32) {
  if (fp32)
    hipLaunchKernelGGL((k_batch_micro_sweep<float4, float>), dim3(nblocks), dim3(HB_BLK), 0, st, probs,
                       bmap, parity);
  else
    hipLaunchKernelGGL((k_batch_micro_sweep<double2, double>), dim3(nblocks), dim3(HB_BLK), 0, st,
                       probs, bmap, parity);
}

int sa_block_size() { return HB_BLK; }
int kp_gauss_per_site() { return KP_NG; }
int kp_unif_per_site() { return KP_NU; }

}  // namespace gfx
