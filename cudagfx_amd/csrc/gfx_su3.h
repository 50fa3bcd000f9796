// gfx_su3.h -- SU(3) algebra for the gfx950 kernels.
//
// Complex 3x3 matrices live entirely in VGPRs (every loop is unrolled, no
// dynamic indexing).  The operation order of each routine follows the
// reference device code (dev_su3.cu, overrelax.cu) so that the FP64 path
// reproduces its arithmetic up to fused multiply-add rounding:
//   * products accumulate k = 0,1,2 as  s = v[i][k]*w[k][j] + s
//     (dev_su3.cu:1418-1475);
//   * the third row of a 2-row matrix is conj(row0 x row1)
//     (dev_su3.cu:289-305, 1617-1632);
//   * Gram-Schmidt normalisation with rsqrt (dev_su3.cu:1569-1634).
// Templated on the arithmetic type R (double for the plain path, float for
// the bulk sweeps of the mixed path).
#pragma once

#include <hip/hip_runtime.h>

namespace gfx {

template <typename R>
struct Cx {
  R re, im;
};

template <typename R>
__device__ __forceinline__ Cx<R> cx(R re, R im) {
  Cx<R> c;
  c.re = re;
  c.im = im;
  return c;
}
template <typename R>
__device__ __forceinline__ Cx<R> cmul(Cx<R> a, Cx<R> b) {
  return cx<R>(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
// a * conj(b)
template <typename R>
__device__ __forceinline__ Cx<R> cmulc(Cx<R> a, Cx<R> b) {
  return cx<R>(a.re * b.re + a.im * b.im, a.im * b.re - a.re * b.im);
}
// Products accumulate through explicit fused multiply-adds (one instruction
// each, the sign of an operand free): a complex multiply-accumulate is four
// of them instead of two multiplies, two multiply-adds and two adds.  The
// k-order of every sum is the reference's; the roundings are those of the
// fused operations.
template <typename R>
__device__ __forceinline__ R fmaR(R a, R b, R c);
template <>
__device__ __forceinline__ double fmaR<double>(double a, double b, double c) {
  return __builtin_fma(a, b, c);
}
template <>
__device__ __forceinline__ float fmaR<float>(float a, float b, float c) {
  return __builtin_fmaf(a, b, c);
}
// s += a b
template <typename R>
__device__ __forceinline__ void cmac(Cx<R>& s, Cx<R> a, Cx<R> b) {
  s.re = fmaR(a.re, b.re, s.re);
  s.re = fmaR(-a.im, b.im, s.re);
  s.im = fmaR(a.re, b.im, s.im);
  s.im = fmaR(a.im, b.re, s.im);
}
// s += a conj(b)
template <typename R>
__device__ __forceinline__ void cmacc(Cx<R>& s, Cx<R> a, Cx<R> b) {
  s.re = fmaR(a.re, b.re, s.re);
  s.re = fmaR(a.im, b.im, s.re);
  s.im = fmaR(a.im, b.re, s.im);
  s.im = fmaR(-a.re, b.im, s.im);
}
// s -= a b
template <typename R>
__device__ __forceinline__ void cmsb(Cx<R>& s, Cx<R> a, Cx<R> b) {
  s.re = fmaR(-a.re, b.re, s.re);
  s.re = fmaR(a.im, b.im, s.re);
  s.im = fmaR(-a.re, b.im, s.im);
  s.im = fmaR(-a.im, b.re, s.im);
}
// conj(a b - c d)
template <typename R>
__device__ __forceinline__ Cx<R> cross_conj(Cx<R> a, Cx<R> b, Cx<R> c, Cx<R> d) {
  Cx<R> t = cmul(a, b);
  cmsb(t, c, d);
  return cconj(t);
}

template <typename R>
__device__ __forceinline__ Cx<R> cadd(Cx<R> a, Cx<R> b) {
  return cx<R>(a.re + b.re, a.im + b.im);
}
template <typename R>
__device__ __forceinline__ Cx<R> csub(Cx<R> a, Cx<R> b) {
  return cx<R>(a.re - b.re, a.im - b.im);
}
template <typename R>
__device__ __forceinline__ Cx<R> cconj(Cx<R> a) {
  return cx<R>(a.re, -a.im);
}

template <typename R>
struct M3 {
  Cx<R> e[3][3];
};

template <typename R>
__device__ __forceinline__ void m_zero(M3<R>& m) {
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) m.e[i][j] = cx<R>(R(0), R(0));
}

// F (default): products accumulated through the fused multiply-adds above;
// F = false: the separate products and sums of the earlier form of these
// routines, kept for the link heatbath and the Haar-random g, so that the
// configurations and random starts they generate stay what they were.

// third row = conj(row0 x row1)
template <typename R, bool F = true>
__device__ __forceinline__ void m_row2(M3<R>& u) {
  if constexpr (F) {
    u.e[2][0] = cross_conj(u.e[0][1], u.e[1][2], u.e[0][2], u.e[1][1]);
    u.e[2][1] = cross_conj(u.e[0][2], u.e[1][0], u.e[0][0], u.e[1][2]);
    u.e[2][2] = cross_conj(u.e[0][0], u.e[1][1], u.e[0][1], u.e[1][0]);
  } else {
    u.e[2][0] = cconj(csub(cmul(u.e[0][1], u.e[1][2]), cmul(u.e[0][2], u.e[1][1])));
    u.e[2][1] = cconj(csub(cmul(u.e[0][2], u.e[1][0]), cmul(u.e[0][0], u.e[1][2])));
    u.e[2][2] = cconj(csub(cmul(u.e[0][0], u.e[1][1]), cmul(u.e[0][1], u.e[1][0])));
  }
}

// The same third row without fused multiply-adds: bit for bit what the
// reference's host conversion from2vtosu3_trafo (su3manip.c:341-356)
// computes from the stored rows.  Used where matrices go back to the host.
template <typename R>
__device__ __forceinline__ void m_row2_unfused(M3<R>& u) {
#pragma clang fp contract(off)
  const int ia[3] = {1, 2, 0}, ib[3] = {2, 0, 1};
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const Cx<R> a1 = u.e[0][ia[c]], b2 = u.e[1][ib[c]];
    const Cx<R> a2 = u.e[0][ib[c]], b1 = u.e[1][ia[c]];
    const R h1re = a1.re * b2.re - a1.im * b2.im;
    const R h1im = a1.re * b2.im + a1.im * b2.re;
    const R h2re = a2.re * b1.re - a2.im * b1.im;
    const R h2im = a2.re * b1.im + a2.im * b1.re;
    u.e[2][c].re = h1re - h2re;
    u.e[2][c].im = R(-1) * (h1im - h2im);
  }
}

// Dagger of a 2-row matrix whose rows sit in u.e[0..1]: the result's
// columns 0,1 are the conjugated rows, column 2 the conjugated cross product
// (dev_su3.cu:826-929).
template <typename R, bool F = true>
__device__ __forceinline__ void m_dagger_from_rows(M3<R>& d, const Cx<R> r0[3], const Cx<R> r1[3]) {
#pragma unroll
  for (int i = 0; i < 3; i++) {
    d.e[i][0] = cconj(r0[i]);
    d.e[i][1] = cconj(r1[i]);
  }
  if constexpr (F) {
    d.e[0][2] = cross_conj(d.e[1][0], d.e[2][1], d.e[2][0], d.e[1][1]);
    d.e[1][2] = cross_conj(d.e[2][0], d.e[0][1], d.e[0][0], d.e[2][1]);
    d.e[2][2] = cross_conj(d.e[0][0], d.e[1][1], d.e[1][0], d.e[0][1]);
  } else {
    d.e[0][2] = csub(cconj(cmul(d.e[1][0], d.e[2][1])), cconj(cmul(d.e[2][0], d.e[1][1])));
    d.e[1][2] = csub(cconj(cmul(d.e[2][0], d.e[0][1])), cconj(cmul(d.e[0][0], d.e[2][1])));
    d.e[2][2] = csub(cconj(cmul(d.e[0][0], d.e[1][1])), cconj(cmul(d.e[1][0], d.e[0][1])));
  }
}

// u += v * w   (dev_add_su3_ti_su3)
template <typename R, bool F = true>
__device__ __forceinline__ void m_addmul(M3<R>& u, const M3<R>& v, const M3<R>& w) {
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      if constexpr (F) {
        cmac(u.e[i][j], v.e[i][0], w.e[0][j]);
        cmac(u.e[i][j], v.e[i][1], w.e[1][j]);
        cmac(u.e[i][j], v.e[i][2], w.e[2][j]);
      } else {
        Cx<R> s = cmul(v.e[i][0], w.e[0][j]);
        s = cadd(cmul(v.e[i][1], w.e[1][j]), s);
        s = cadd(cmul(v.e[i][2], w.e[2][j]), s);
        u.e[i][j].re += s.re;
        u.e[i][j].im += s.im;
      }
    }
}

// u = v * w   (dev_su3_ti_su3)
template <typename R, bool F = true>
__device__ __forceinline__ M3<R> m_mul(const M3<R>& v, const M3<R>& w) {
  M3<R> u;
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      Cx<R> s = cmul(v.e[i][0], w.e[0][j]);
      if constexpr (F) {
        cmac(s, v.e[i][1], w.e[1][j]);
        cmac(s, v.e[i][2], w.e[2][j]);
      } else {
        s = cadd(cmul(v.e[i][1], w.e[1][j]), s);
        s = cadd(cmul(v.e[i][2], w.e[2][j]), s);
      }
      u.e[i][j] = s;
    }
  return u;
}

// u = v * w^+   (dev_su3_ti_su3d)
template <typename R>
__device__ __forceinline__ M3<R> m_mul_d(const M3<R>& v, const M3<R>& w) {
  M3<R> u;
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      Cx<R> s = cmulc(v.e[i][0], w.e[j][0]);
      cmacc(s, v.e[i][1], w.e[j][1]);
      cmacc(s, v.e[i][2], w.e[j][2]);
      u.e[i][j] = s;
    }
  return u;
}

// element (i,j) of v * w
template <typename R, bool F = true>
__device__ __forceinline__ Cx<R> m_mul_elem(const M3<R>& v, const M3<R>& w, int i, int j) {
  Cx<R> s = cmul(v.e[i][0], w.e[0][j]);
  if constexpr (F) {
    cmac(s, v.e[i][1], w.e[1][j]);
    cmac(s, v.e[i][2], w.e[2][j]);
  } else {
    s = cadd(cmul(v.e[i][1], w.e[1][j]), s);
    s = cadd(cmul(v.e[i][2], w.e[2][j]), s);
  }
  return s;
}

template <typename R>
__device__ __forceinline__ R rsqrt_t(R x);
template <>
__device__ __forceinline__ double rsqrt_t<double>(double x) {
  return ::rsqrt(x);
}
template <>
__device__ __forceinline__ float rsqrt_t<float>(float x) {
  return ::rsqrtf(x);
}

// Gram-Schmidt on rows 0,1; row 2 from the cross product (dev_su3.cu:1569-1634)
template <typename R, bool F = true>
__device__ __forceinline__ void m_normalize(M3<R>& u) {
  R len = (u.e[0][0].re * u.e[0][0].re + u.e[0][0].im * u.e[0][0].im) +
          (u.e[0][1].re * u.e[0][1].re + u.e[0][1].im * u.e[0][1].im) +
          (u.e[0][2].re * u.e[0][2].re + u.e[0][2].im * u.e[0][2].im);
  len = rsqrt_t<R>(len);
#pragma unroll
  for (int i = 0; i < 3; i++) {
    u.e[0][i].re = u.e[0][i].re * len;
    u.e[0][i].im = u.e[0][i].im * len;
  }
  Cx<R> proj = cmulc(u.e[1][0], u.e[0][0]);
  if constexpr (F) {
    cmacc(proj, u.e[1][1], u.e[0][1]);
    cmacc(proj, u.e[1][2], u.e[0][2]);
#pragma unroll
    for (int i = 0; i < 3; i++) cmsb(u.e[1][i], proj, u.e[0][i]);
  } else {
    proj = cadd(proj, cmulc(u.e[1][1], u.e[0][1]));
    proj = cadd(proj, cmulc(u.e[1][2], u.e[0][2]));
#pragma unroll
    for (int i = 0; i < 3; i++) u.e[1][i] = csub(u.e[1][i], cmul(proj, u.e[0][i]));
  }
  len = (u.e[1][0].re * u.e[1][0].re + u.e[1][0].im * u.e[1][0].im) +
        (u.e[1][1].re * u.e[1][1].re + u.e[1][1].im * u.e[1][1].im) +
        (u.e[1][2].re * u.e[1][2].re + u.e[1][2].im * u.e[1][2].im);
  len = rsqrt_t<R>(len);
#pragma unroll
  for (int i = 0; i < 3; i++) {
    u.e[1][i].re = u.e[1][i].re * len;
    u.e[1][i].im = u.e[1][i].im * len;
  }
  m_row2<R, F>(u);
}

// Left multiplication of rows a,b of g by the SU(2) element
// [[w0 + i w3, w2 + i w1], [-w2 + i w1, w0 - i w3]]  (overrelax.cu:244-265)
template <typename R, bool F = true>
__device__ __forceinline__ void su2_left(M3<R>& g, int a, int b, R w0, R w1, R w2, R w3) {
#pragma unroll
  for (int c = 0; c < 3; c++) {
    if constexpr (F) {
      Cx<R> d = cmul(cx<R>(w0, w3), g.e[a][c]);
      cmac(d, cx<R>(w2, w1), g.e[b][c]);
      Cx<R> e = cmul(cx<R>(-w2, w1), g.e[a][c]);
      cmac(e, cx<R>(w0, -w3), g.e[b][c]);
      g.e[b][c] = e;
      g.e[a][c] = d;
    } else {
      Cx<R> d = cadd(cmul(cx<R>(w0, w3), g.e[a][c]), cmul(cx<R>(w2, w1), g.e[b][c]));
      g.e[b][c] = cadd(cmul(cx<R>(-w2, w1), g.e[a][c]), cmul(cx<R>(w0, -w3), g.e[b][c]));
      g.e[a][c] = d;
    }
  }
}

// The SU(2) subgroup vector w of X = g*K for rows (a,b)  (overrelax.cu:217-223).
// Only the four entries of X that enter are formed; each with the same
// k-order as the full product, so the values equal those of the full product.
template <typename R, bool F = true>
__device__ __forceinline__ void su2_subvec(const M3<R>& g, const M3<R>& K, int a, int b,
                                           R& w0, R& w1, R& w2, R& w3) {
  Cx<R> xaa = m_mul_elem<R, F>(g, K, a, a);
  Cx<R> xbb = m_mul_elem<R, F>(g, K, b, b);
  Cx<R> xab = m_mul_elem<R, F>(g, K, a, b);
  Cx<R> xba = m_mul_elem<R, F>(g, K, b, a);
  w0 = xaa.re + xbb.re;
  w3 = -xaa.im + xbb.im;
  w1 = -xab.im - xba.im;
  w2 = -xab.re + xba.re;
}

// Cabibbo-Marinari maximisation of Re tr(g K) over the subgroups
// (0,1),(0,2),(1,2) (overrelax.cu:188-269)
template <typename R>
__device__ __forceinline__ void cm_relax(M3<R>& g, const M3<R>& K) {
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = a + 1; b < 3; b++) {
      R w0, w1, w2, w3;
      su2_subvec(g, K, a, b, w0, w1, w2, w3);
      R det = w0 * w0 + w1 * w1 + w2 * w2 + w3 * w3;
      det = rsqrt_t<R>(det);
      su2_left(g, a, b, w0 * det, w1 * det, w2 * det, w3 * det);
    }
}

// g <- normalize(1 + w R + w(w-1)/2 R^2) g with R = gn g^+ - 1: the
// reference's 3-term binomial series (overrelax.cu:149-177, N = 3).
// c1 = (1+w-1)/1 and c2 = (1+w-2)/2 are formed on the host exactly as the
// reference forms them.  For general N the loop form is kept.
template <typename R>
__device__ __forceinline__ void overrelax_apply(M3<R>& g, const M3<R>& gn, R c1, R c2) {
  M3<R> Rm = m_mul_d(gn, g);
  Rm.e[0][0].re -= R(1);
  Rm.e[1][1].re -= R(1);
  Rm.e[2][2].re -= R(1);
  // a1 = (1 * R) * c1 ; Rw = 1 + a1
  M3<R> a1, Rw;
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      a1.e[i][j].re = Rm.e[i][j].re * c1;
      a1.e[i][j].im = Rm.e[i][j].im * c1;
      Rw.e[i][j] = a1.e[i][j];
    }
  Rw.e[0][0].re = R(1) + a1.e[0][0].re;
  Rw.e[1][1].re = R(1) + a1.e[1][1].re;
  Rw.e[2][2].re = R(1) + a1.e[2][2].re;
  // a2 = (a1 * R) * c2 ; Rw = Rw + a2
  M3<R> a2 = m_mul(a1, Rm);
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      Rw.e[i][j].re = Rw.e[i][j].re + a2.e[i][j].re * c2;
      Rw.e[i][j].im = Rw.e[i][j].im + a2.e[i][j].im * c2;
    }
  m_normalize(Rw);
  g = m_mul(Rw, g);
}

// ---- shorter arithmetic for the FP32 sweep (the FP64 sweep keeps the
// reference's operation order).  Both use that a product of SU(3) matrices
// is in SU(3): two rows are formed and the third is conj(row0 x row1).

// a[0..1] are rows 0,1 of A, b[0..2] the rows of B: rows 0,1 of A B^+
template <typename R>
__device__ __forceinline__ void rows_mul_dagger(Cx<R> p[2][3], const Cx<R> a[2][3],
                                                const Cx<R> b[3][3]) {
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      p[i][j] = cmulc(a[i][0], b[j][0]);
      cmacc(p[i][j], a[i][1], b[j][1]);
      cmacc(p[i][j], a[i][2], b[j][2]);
    }
}
// rows 0,1 of A B
template <typename R>
__device__ __forceinline__ void rows_mul(Cx<R> p[2][3], const Cx<R> a[2][3], const Cx<R> b[3][3]) {
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      p[i][j] = cmul(a[i][0], b[0][j]);
      cmac(p[i][j], a[i][1], b[1][j]);
      cmac(p[i][j], a[i][2], b[2][j]);
    }
}
// r2 = conj(r0 x r1)
template <typename R>
__device__ __forceinline__ void row2_of(Cx<R> r2[3], const Cx<R> r0[3], const Cx<R> r1[3]) {
  r2[0] = cross_conj(r0[1], r1[2], r0[2], r1[1]);
  r2[1] = cross_conj(r0[2], r1[0], r0[0], r1[2]);
  r2[2] = cross_conj(r0[0], r1[1], r0[1], r1[0]);
}

// g <- normalize(Rw g) with Rw = 1 + c1 R + c1 c2 R^2, R = h g^+ - 1, in the
// form Rw g = g + c1 (h - g) + c1 c2 (h g^+ h - 2 h + g): two 2x3 products
// instead of three 3x3 ones.  Gram-Schmidt commutes with multiplying by the
// unitary g from the right, so normalize(Rw) g = normalize(Rw g).
template <typename R>
__device__ __forceinline__ void overrelax_apply_short(M3<R>& g, const M3<R>& h, R c1, R c2) {
  Cx<R> t[2][3], w[2][3];
  rows_mul_dagger(t, h.e, g.e);
  rows_mul(w, t, h.e);
  const R cw = c1 * c2;
  const R cg = R(1) - c1 + cw, ch = c1 - R(2) * cw;
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      g.e[i][j].re = cg * g.e[i][j].re + ch * h.e[i][j].re + cw * w[i][j].re;
      g.e[i][j].im = cg * g.e[i][j].im + ch * h.e[i][j].im + cw * w[i][j].im;
    }
  m_normalize(g);
}

// traceless anti-hermitian part (i/2)(U^+ - U) - tr/3 (overrelax.cu:451-473)
template <typename R>
__device__ __forceinline__ void vecpot(M3<R>& U) {
  M3<R> t;
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      Cx<R> d = csub(cconj(U.e[j][i]), U.e[i][j]);
      // (0, 0.5) * d
      t.e[i][j] = cx<R>(R(0) * d.re - R(0.5) * d.im, R(0) * d.im + R(0.5) * d.re);
    }
  Cx<R> tr = cadd(cadd(t.e[0][0], t.e[1][1]), t.e[2][2]);
  tr.re = tr.re / R(3);
  tr.im = tr.im / R(3);
#pragma unroll
  for (int i = 0; i < 3; i++) t.e[i][i] = csub(t.e[i][i], tr);
  U = t;
}

template <typename R>
__device__ __forceinline__ R m_retrace(const M3<R>& m) {
  return (m.e[0][0].re + m.e[1][1].re) + m.e[2][2].re;
}

}  // namespace gfx
