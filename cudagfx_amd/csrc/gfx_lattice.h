// gfx_lattice.h -- checkerboard geometry and the structure-of-arrays layout.
//
// Sites: lexicographic s = x0 + L0*(x1 + L1*(x2 + L2*x3)) as in the
// reference (cudagaugefix.cu:79-85).  Parity p = (x0+x1+x2+x3) & 1.  With L0
// even, the index of a site inside its parity is i = s >> 1, which is the
// reference's even/odd order (lexicographic filtered by parity,
// cudagaugefix.cu:159-195), so eoidx_{even,odd}[i] == 2*i + (bit of x0).
//
// Device layout.  A *field* holds one 2-row SU(3) matrix (rows 0,1 = 12
// reals, su3manip.c:212-230) for each of the nh = V/2 sites of one parity,
// as planes of 16-byte vectors so that a wavefront reads 64 consecutive
// vectors (1 KiB) per load:
//   FP64: 6 planes of double2  = (u00)(u01)(u02)(u10)(u11)(u12)
//   FP32: 3 planes of float4   = (u00,u01)(u02,u10)(u11,u12)
// plane c of site i sits at field[c*stride + i]; stride >= nh, a multiple
// of 64.  Links: field (p*4 + mu) holds U_mu(x) for x of parity p.  Gauge
// transformation g: field p.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "gfx_su3.h"

namespace gfx {

// n / d for 0 <= n < 2^31 with a precomputed multiplier ("round-up" method
// of Granlund & Montgomery).
struct FastDiv {
  uint32_t d, m, s;
};

static inline FastDiv make_fastdiv(uint32_t d) {
  FastDiv f;
  f.d = d;
  uint32_t s = 0;
  while ((1ull << s) < d) s++;
  f.s = s;
  f.m = (uint32_t)((((1ull << 32) * ((1ull << s) - d)) / d) + 1);
  return f;
}

__host__ __device__ __forceinline__ uint32_t fdiv(uint32_t n, FastDiv f) {
#ifdef __HIP_DEVICE_COMPILE__
  uint32_t t = __umulhi(n, f.m);
#else
  uint32_t t = (uint32_t)(((uint64_t)n * f.m) >> 32);
#endif
  return (t + n) >> f.s;
}

struct Geom {
  int L[4];      // extents x0..x3
  int h0;        // L0/2
  int nh;        // sites per parity
  int st[4];     // index strides inside a parity: 1 (x0 special), h0, h0*L1, h0*L1*L2
  int stride;    // plane stride (elements)
  FastDiv dh0, dL1, dL2;
};

struct Site {
  int x[4];
  int q;  // i / h0
};

__device__ __forceinline__ Site site_coords(const Geom& g, int i, int parity) {
  Site s;
  uint32_t q = fdiv((uint32_t)i, g.dh0);
  int i0h = i - (int)q * g.h0;
  uint32_t q2 = fdiv(q, g.dL1);
  s.x[1] = (int)(q - q2 * (uint32_t)g.L[1]);
  uint32_t q3 = fdiv(q2, g.dL2);
  s.x[2] = (int)(q2 - q3 * (uint32_t)g.L[2]);
  s.x[3] = (int)q3;
  s.x[0] = 2 * i0h + ((parity + s.x[1] + s.x[2] + s.x[3]) & 1);
  s.q = (int)q;
  return s;
}

// index (in the other parity) of the forward / backward neighbour in mu
__device__ __forceinline__ int nb_fwd(const Geom& g, const Site& s, int i, int mu) {
  if (mu == 0) {
    int x0 = s.x[0] + 1;
    if (x0 == g.L[0]) x0 = 0;
    return (x0 >> 1) + s.q * g.h0;
  }
  return (s.x[mu] == g.L[mu] - 1) ? i - (g.L[mu] - 1) * g.st[mu] : i + g.st[mu];
}
__device__ __forceinline__ int nb_bwd(const Geom& g, const Site& s, int i, int mu) {
  if (mu == 0) {
    int x0 = (s.x[0] == 0) ? g.L[0] - 1 : s.x[0] - 1;
    return (x0 >> 1) + s.q * g.h0;
  }
  return (s.x[mu] == 0) ? i + (g.L[mu] - 1) * g.st[mu] : i - g.st[mu];
}

// ---- storage traits: 2-row matrices in planes of 16-byte vectors ----
// NT = true issues non-temporal (streaming) loads: used for the links, which
// a half sweep reads exactly once, so that they do not evict the neighbour g
// lines that other sites still need from L2.
typedef double nat_d2 __attribute__((ext_vector_type(2)));
typedef float nat_f4 __attribute__((ext_vector_type(4)));

// element i of a plane: the plane's start (uniform across a wavefront) plus
// a 32-bit byte offset, so that loads and stores take the scalar-base form
// with one 32-bit lane offset instead of a 64-bit address per lane.
// Requires i * sizeof(T) < 2^32: at most 2^28 sites per parity (geometry_ok).
template <typename T>
__device__ __forceinline__ const T* at(const T* plane, int i) {
  return reinterpret_cast<const T*>(reinterpret_cast<const char*>(plane) +
                                    (uint32_t)i * (uint32_t)sizeof(T));
}
template <typename T>
__device__ __forceinline__ T* at(T* plane, int i) {
  return reinterpret_cast<T*>(reinterpret_cast<char*>(plane) + (uint32_t)i * (uint32_t)sizeof(T));
}

template <bool NT>
__device__ __forceinline__ double2 ld16(const double2* p) {
  if (NT) {
    const nat_d2 v = __builtin_nontemporal_load(reinterpret_cast<const nat_d2*>(p));
    return make_double2(v.x, v.y);
  }
  return *p;
}
template <bool NT>
__device__ __forceinline__ float4 ld16(const float4* p) {
  if (NT) {
    const nat_f4 v = __builtin_nontemporal_load(reinterpret_cast<const nat_f4*>(p));
    return make_float4(v.x, v.y, v.z, v.w);
  }
  return *p;
}

template <typename V>
struct Store;

template <>
struct Store<double2> {
  static constexpr int NPLANE = 6;
  template <typename R, bool NT = false>
  __device__ __forceinline__ static void load(const double2* __restrict__ f, int stride, int i,
                                              Cx<R> r0[3], Cx<R> r1[3]) {
    double2 v[6];
#pragma unroll
    for (int c = 0; c < 6; c++) v[c] = ld16<NT>(at(f + (size_t)c * stride, i));
#pragma unroll
    for (int c = 0; c < 3; c++) {
      r0[c] = cx<R>((R)v[c].x, (R)v[c].y);
      r1[c] = cx<R>((R)v[3 + c].x, (R)v[3 + c].y);
    }
  }
  template <typename R>
  __device__ __forceinline__ static void store(double2* __restrict__ f, int stride, int i,
                                               const M3<R>& m) {
#pragma unroll
    for (int c = 0; c < 3; c++) {
      *at(f + (size_t)c * stride, i) = make_double2((double)m.e[0][c].re, (double)m.e[0][c].im);
      *at(f + (size_t)(3 + c) * stride, i) =
          make_double2((double)m.e[1][c].re, (double)m.e[1][c].im);
    }
  }
};

template <>
struct Store<float4> {
  static constexpr int NPLANE = 3;
  template <typename R, bool NT = false>
  __device__ __forceinline__ static void load(const float4* __restrict__ f, int stride, int i,
                                              Cx<R> r0[3], Cx<R> r1[3]) {
    float4 a = ld16<NT>(at(f, i));
    float4 b = ld16<NT>(at(f + (size_t)stride, i));
    float4 c = ld16<NT>(at(f + (size_t)2 * stride, i));
    r0[0] = cx<R>((R)a.x, (R)a.y);
    r0[1] = cx<R>((R)a.z, (R)a.w);
    r0[2] = cx<R>((R)b.x, (R)b.y);
    r1[0] = cx<R>((R)b.z, (R)b.w);
    r1[1] = cx<R>((R)c.x, (R)c.y);
    r1[2] = cx<R>((R)c.z, (R)c.w);
  }
  template <typename R>
  __device__ __forceinline__ static void store(float4* __restrict__ f, int stride, int i,
                                               const M3<R>& m) {
    *at(f, i) = make_float4((float)m.e[0][0].re, (float)m.e[0][0].im, (float)m.e[0][1].re,
                       (float)m.e[0][1].im);
    *at(f + (size_t)stride, i) = make_float4((float)m.e[0][2].re, (float)m.e[0][2].im,
                                        (float)m.e[1][0].re, (float)m.e[1][0].im);
    *at(f + (size_t)2 * stride, i) = make_float4((float)m.e[1][1].re, (float)m.e[1][1].im,
                                            (float)m.e[1][2].re, (float)m.e[1][2].im);
  }
};

// ---- 32-bit fixed-point fields (links and g of the mixed Landau run's
// FP64 polish) ----
// The elements u of a unitary matrix lie in [-1, 1]: each real is held as
// the int32 q = rint(u 2^31), clamped to +-(2^31 - 1), in 3 planes of int4
// in the order of the FP32 planes, (u00,u01)(u02,u10)(u11,u12) as (re, im)
// pairs, 48 bytes per matrix.  q 2^-31 is within 2^-32 of u (2^-31 where
// |u| rounds to 1 and q is clamped).  The arithmetic on them is FP64.
constexpr double FX_SCALE = 0x1p31;
// A polish on fixed-point links and g is used for stop bounds eps above
// this: the rounding of g after every update and of the links (2^-32 per
// real) moves Delta by about 1e-9, so max dAdA cannot go much below 1e-17.
constexpr double FX_EPS_MIN = 1e-15;
// Worst effect of the fixed-point rounding on a check's Delta = -A(G K)
// when the result is g' = normalize(G) on the exact links (Coulomb slices,
// whose stops are decided by the fused checks): the links contribute 83.2 e
// (as for the packed links, e = 2^-32), and |g' - G|_F <= 31.2 e (G is
// within 10.4 e of the unitary matrix it was rounded from, which
// normalisation moves by at most twice that) enters G K once through G and
// once through the 8 star terms, 16 x 31.2 e: 582 e, taken as 600 e.
constexpr double FX_DDELTA = 600.0 * 0x1p-32;
// runs whose stops are decided with that bound (Coulomb slices, batch
// lattices) use the fixed-point polish only for eps above this: the bound
// then takes at most a quarter of sqrt(eps) (below it, sqrt(eps) would
// approach FX_DDELTA and no check could stop)
constexpr double FX_EPS_BOUND_MIN = 16.0 * FX_DDELTA * FX_DDELTA;
__device__ __forceinline__ int fx_of(double u) {
  return (int)fmin(fmax(rint(u * FX_SCALE), -2147483647.0), 2147483647.0);
}
typedef int nat_i4 __attribute__((ext_vector_type(4)));
template <bool NT>
__device__ __forceinline__ int4 ld16(const int4* p) {
  if (NT) {
    const nat_i4 v = __builtin_nontemporal_load(reinterpret_cast<const nat_i4*>(p));
    return make_int4(v.x, v.y, v.z, v.w);
  }
  return *p;
}
template <typename R>
__device__ __forceinline__ void fx_rows(const int4& a, const int4& b, const int4& c, Cx<R> r0[3],
                                        Cx<R> r1[3]) {
  const R sc = R(1.0 / FX_SCALE);
  r0[0] = cx<R>((R)a.x * sc, (R)a.y * sc);
  r0[1] = cx<R>((R)a.z * sc, (R)a.w * sc);
  r0[2] = cx<R>((R)b.x * sc, (R)b.y * sc);
  r1[0] = cx<R>((R)b.z * sc, (R)b.w * sc);
  r1[1] = cx<R>((R)c.x * sc, (R)c.y * sc);
  r1[2] = cx<R>((R)c.z * sc, (R)c.w * sc);
}

template <>
struct Store<int4> {
  static constexpr int NPLANE = 3;
  template <typename R, bool NT = false>
  __device__ __forceinline__ static void load(const int4* __restrict__ f, int stride, int i,
                                              Cx<R> r0[3], Cx<R> r1[3]) {
    const int4 a = ld16<NT>(at(f, i));
    const int4 b = ld16<NT>(at(f + (size_t)stride, i));
    const int4 c = ld16<NT>(at(f + (size_t)2 * stride, i));
    fx_rows<R>(a, b, c, r0, r1);
  }
  template <typename R>
  __device__ __forceinline__ static void store(int4* __restrict__ f, int stride, int i,
                                               const M3<R>& m) {
    *at(f, i) = make_int4(fx_of((double)m.e[0][0].re), fx_of((double)m.e[0][0].im),
                               fx_of((double)m.e[0][1].re), fx_of((double)m.e[0][1].im));
    *at(f + (size_t)stride, i) =
        make_int4(fx_of((double)m.e[0][2].re), fx_of((double)m.e[0][2].im),
                  fx_of((double)m.e[1][0].re), fx_of((double)m.e[1][0].im));
    *at(f + (size_t)2 * stride, i) =
        make_int4(fx_of((double)m.e[1][1].re), fx_of((double)m.e[1][1].im),
                  fx_of((double)m.e[1][2].re), fx_of((double)m.e[1][2].im));
  }
};

template <typename V, typename R, bool NT = false, bool F = true>
__device__ __forceinline__ M3<R> load_m(const V* f, int stride, int i) {
  M3<R> m;
  Store<V>::template load<R, NT>(f, stride, i, m.e[0], m.e[1]);
  m_row2<R, F>(m);
  return m;
}

template <typename V, typename R, bool NT = false, bool F = true>
__device__ __forceinline__ M3<R> load_m_dagger(const V* f, int stride, int i) {
  Cx<R> r0[3], r1[3];
  Store<V>::template load<R, NT>(f, stride, i, r0, r1);
  M3<R> d;
  m_dagger_from_rows<R, F>(d, r0, r1);
  return d;
}

template <typename V>
__device__ __forceinline__ const V* field_ptr(const V* base, int field, int stride) {
  return base + (size_t)field * Store<V>::NPLANE * stride;
}
template <typename V>
__device__ __forceinline__ V* field_ptr(V* base, int field, int stride) {
  return base + (size_t)field * Store<V>::NPLANE * stride;
}

// ---- packed links of the mixed paths ----
// The elements u of a unitary link lie in [-1, 1], so they are held in fixed
// point.  For the FP32 bulk sweeps: q = rint(u 2^20), clamped to
// +-(2^20 - 1), a signed 21-bit integer; the 12 reals of a link take 252 of
// the 256 bits of two uint4 planes (a21: dwords 0-3, b21: dwords 4-7), real
// k at bits [21k, 21k + 21).  They read q 2^-20, within 2^-21 of u (2^-20
// where |u| rounds to 1 and q is clamped): 32 instead of 48 bytes per link,
// in 16-byte loads only.  That rounding moves a site's Delta by at most
// 83.2 2^-21 = 4e-5 against the 1e-3 of sqrt(theta) at which the FP32 phase
// ends by default; its checks decide no stop.
// For the FP64 polish on packed links (Coulomb and batch runs below the
// fixed-point polish's reach): hi = rint(u 2^23), clamped to +-(2^23 - 1), a
// signed 24-bit integer; the 12 reals of a link take 288 bits = 9 dwords,
// stored as two uint4 planes (dwords 0-3, 4-7) and one uint plane (dword 8),
// real k at bits [24k, 24k + 24); hi 2^-23 is within 2^-24 of u (2^-23 where
// |u| rounds to 1 and hi is clamped).  The FP64 polish adds lo = rint((u - hi 2^-23) 2^38), an int16
// (|u - hi 2^-23| <= 2^-23 even where hi is clamped, so |lo| <= 2^15,
// clamped to 32767): hi 2^-23 + lo 2^-38 is within 2^-39 of u (2^-38 where
// both clamp), 60 instead of 96 bytes per link.  lo is held in three uint2 planes per
// field, four int16 per 8-byte word in the order of the reals.
constexpr double PACK_HI = 0x1p23;
constexpr double PACK21 = 0x1p20;
constexpr double PACK_LO = 0x1p38;
// Worst effect of the packing on a site's Delta = -A(g K) (A: traceless
// antihermitian part, which does not increase the Frobenius norm).  A stored
// real is off by at most e (2^-38 for high + low parts, where |u| rounds to
// 1); row 2 = conj(row0 x row1) then by at most 4 sqrt(2) e per complex
// element, so |dU|_F <= sqrt(12 + 96) e <= 10.4 e, and with 8 star terms of
// unitary g, |dDelta|_F <= 83.2 e.  theta(x) = |Delta(x)|_F^2, so theta on
// the exact links is below (sqrt(theta on the packed links) + PACK_DDELTA)^2.
// The short form of the star (star_terms_short: rows 0, 1 of each product,
// row 2 rebuilt) adds the rebuild of the product's row 2: rows 0, 1 of U g^+
// are off by at most sqrt(6) e per complex element, its row 2 by 4 sqrt(6)
// e, so a term by sqrt(12 + 288) e <= 17.4 e and Delta by 139.2 e
// (PACK_DDELTA64_SHORT).  These bound the FP64 checks, whose own rounding
// (about 1e-15 in Delta) is far below them.  Checks made in FP32 (FP32 g
// and FP32 arithmetic, rounding of order 1e-6 in Delta) decide no stop: a
// lattice or slice that an FP32 check finds below eps moves to FP64, and an
// FP64 check decides its stop.
constexpr double PACK_DDELTA64 = 83.2 * 0x1p-38;
constexpr double PACK_DDELTA64_SHORT = 139.2 * 0x1p-38;
// runs whose stops are decided with that bound use packed links only for
// eps above this (below it the bound would take most of eps)
constexpr double PACK_EPS_MIN = 100.0 * PACK_DDELTA64_SHORT * PACK_DDELTA64_SHORT;
struct PackedLinks {
  const uint4* a;       // 24-bit hi: field f, site i at [f * stride + i]
  const uint4* b;
  const uint32_t* c;
  const uint2* lo;      // field f, plane p at [(3f + p) * stride + i]; FP64 only
  const uint4* a21;     // 21-bit (FP32 sweeps): field f, site i at [f * stride + i]
  const uint4* b21;
};
// the same planes as the packing kernel writes them (host side)
struct PackedArgs {
  uint4* a = nullptr;
  uint4* b = nullptr;
  uint32_t* c = nullptr;
  uint2* lo = nullptr;
  uint4* a21 = nullptr;
  uint4* b21 = nullptr;
};

typedef unsigned int nat_u2 __attribute__((ext_vector_type(2)));
typedef unsigned int nat_u4 __attribute__((ext_vector_type(4)));
template <bool NT>
__device__ __forceinline__ uint2 ld8(const uint2* p) {
  if (NT) {
    const nat_u2 v = __builtin_nontemporal_load(reinterpret_cast<const nat_u2*>(p));
    return make_uint2(v.x, v.y);
  }
  return *p;
}
template <bool NT>
__device__ __forceinline__ uint4 ld16(const uint4* p) {
  if (NT) {
    const nat_u4 v = __builtin_nontemporal_load(reinterpret_cast<const nat_u4*>(p));
    return make_uint4(v.x, v.y, v.z, v.w);
  }
  return *p;
}
template <bool NT>
__device__ __forceinline__ uint32_t ld4(const uint32_t* p) {
  if (NT) return __builtin_nontemporal_load(p);
  return *p;
}
// signed 24-bit value K of the 288-bit stream w[0..8]
template <int K>
__device__ __forceinline__ int pack_s24(const uint32_t w[9]) {
  constexpr int o = 24 * K, d = o / 32, sh = o % 32;
  uint32_t v;
  if (sh <= 8)
    v = w[d] >> sh;
  else
    v = __builtin_amdgcn_alignbit(w[d + 1], w[d], sh);
  return ((int)(v << 8)) >> 8;
}
__device__ __forceinline__ int pack_lo16(const uint32_t q[6], int k) {
  const uint32_t v = q[k >> 1];
  return (k & 1) ? ((int)v >> 16) : (int)(short)(v & 0xffffu);
}
// signed 21-bit value K of the 256-bit stream w[0..7]
template <int K>
__device__ __forceinline__ int pack_s21(const uint32_t w[8]) {
  constexpr int o = 21 * K, d = o / 32, sh = o % 32;
  if constexpr (sh <= 11) {
    return __builtin_amdgcn_sbfe((int)w[d], sh, 21);
  } else {
    const uint32_t v = __builtin_amdgcn_alignbit(w[d + 1], w[d], sh);
    return __builtin_amdgcn_sbfe((int)v, 0, 21);
  }
}
// the 21-bit q of u, as stored by k_pack_links
__device__ __forceinline__ int pack21_of(double u) {
  return (int)fmin(fmax(rint(u * PACK21), -1048575.0), 1048575.0);
}
// hi and lo of u, as stored by k_pack_links
__device__ __forceinline__ int pack_hi_of(double u) {
  return (int)fmin(fmax(rint(u * PACK_HI), -8388607.0), 8388607.0);
}
__device__ __forceinline__ int pack_lo_of(double u, int hi) {
  return (int)fmin(fmax(rint((u - (double)hi / PACK_HI) * PACK_LO), -32767.0), 32767.0);
}

// A matrix as loaded, before it is turned into rows: the sweeps issue the
// loads of all four operands of a direction (two links, two neighbour g)
// before decoding any of them, so that a wavefront has them in flight
// together instead of one after another.
template <typename LK, typename R>
struct Raw;
template <typename R>
struct Raw<const double2*, R> {
  double2 v[6];
};
template <typename R>
struct Raw<const float4*, R> {
  float4 v[3];
};
template <typename R>
struct Raw<const int4*, R> {
  int4 v[3];
};
template <>
struct Raw<PackedLinks, double> {
  uint4 a, b;
  uint32_t c;
  uint2 lo[3];
};
template <>
struct Raw<PackedLinks, float> {
  uint4 a, b;  // the 21-bit planes
};

template <typename R, bool NT>
__device__ __forceinline__ void raw_load(const double2* f, int stride, int i, Raw<const double2*, R>& r) {
#pragma unroll
  for (int c = 0; c < 6; c++) r.v[c] = ld16<NT>(at(f + (size_t)c * stride, i));
}
template <typename R, bool NT>
__device__ __forceinline__ void raw_load(const float4* f, int stride, int i, Raw<const float4*, R>& r) {
#pragma unroll
  for (int c = 0; c < 3; c++) r.v[c] = ld16<NT>(at(f + (size_t)c * stride, i));
}
template <typename R, bool NT>
__device__ __forceinline__ void raw_load(const int4* f, int stride, int i, Raw<const int4*, R>& r) {
#pragma unroll
  for (int c = 0; c < 3; c++) r.v[c] = ld16<NT>(at(f + (size_t)c * stride, i));
}
template <typename R, bool NT>
__device__ __forceinline__ void raw_load(const PackedLinks& U, int field, int stride, int i,
                                         Raw<PackedLinks, R>& r) {
  const size_t f = (size_t)field * stride;
  if constexpr (std::is_same<R, double>::value) {
    r.a = ld16<NT>(at(U.a + f, i));
    r.b = ld16<NT>(at(U.b + f, i));
    r.c = ld4<NT>(at(U.c + f, i));
    const uint2* l = U.lo + (size_t)field * 3 * stride;
#pragma unroll
    for (int p = 0; p < 3; p++) r.lo[p] = ld8<NT>(at(l + (size_t)p * stride, i));
  } else {
    r.a = ld16<NT>(at(U.a21 + f, i));
    r.b = ld16<NT>(at(U.b21 + f, i));
  }
}
// links given as planes: the field's first plane
template <typename R, bool NT, typename V>
__device__ __forceinline__ void raw_load(const V* U, int field, int stride, int i, Raw<const V*, R>& r) {
  raw_load<R, NT>(field_ptr(U, field, stride), stride, i, r);
}

template <typename R>
__device__ __forceinline__ void raw_rows(const Raw<const double2*, R>& r, Cx<R> r0[3], Cx<R> r1[3]) {
#pragma unroll
  for (int c = 0; c < 3; c++) {
    r0[c] = cx<R>((R)r.v[c].x, (R)r.v[c].y);
    r1[c] = cx<R>((R)r.v[3 + c].x, (R)r.v[3 + c].y);
  }
}
template <typename R>
__device__ __forceinline__ void raw_rows(const Raw<const float4*, R>& r, Cx<R> r0[3], Cx<R> r1[3]) {
  const float4 a = r.v[0], b = r.v[1], c = r.v[2];
  r0[0] = cx<R>((R)a.x, (R)a.y);
  r0[1] = cx<R>((R)a.z, (R)a.w);
  r0[2] = cx<R>((R)b.x, (R)b.y);
  r1[0] = cx<R>((R)b.z, (R)b.w);
  r1[1] = cx<R>((R)c.x, (R)c.y);
  r1[2] = cx<R>((R)c.z, (R)c.w);
}
template <typename R>
__device__ __forceinline__ void raw_rows(const Raw<const int4*, R>& r, Cx<R> r0[3], Cx<R> r1[3]) {
  fx_rows<R>(r.v[0], r.v[1], r.v[2], r0, r1);
}
// R = float: the 21-bit planes; NOSCALE: the integers themselves, 2^20
// times the link (exact in FP32), the star terms of the FP32 sweeps taking
// the scale off where they are added to K.  R = double: hi + lo.
template <typename R, bool NOSCALE = false>
__device__ __forceinline__ void raw_rows(const Raw<PackedLinks, R>& r, Cx<R> r0[3], Cx<R> r1[3]) {
  R v[12];
  if constexpr (std::is_same<R, double>::value) {
    const uint32_t w[9] = {r.a.x, r.a.y, r.a.z, r.a.w, r.b.x, r.b.y, r.b.z, r.b.w, r.c};
    const int h[12] = {pack_s24<0>(w), pack_s24<1>(w), pack_s24<2>(w),  pack_s24<3>(w),
                       pack_s24<4>(w), pack_s24<5>(w), pack_s24<6>(w),  pack_s24<7>(w),
                       pack_s24<8>(w), pack_s24<9>(w), pack_s24<10>(w), pack_s24<11>(w)};
    const uint32_t q[6] = {r.lo[0].x, r.lo[0].y, r.lo[1].x, r.lo[1].y, r.lo[2].x, r.lo[2].y};
#pragma unroll
    for (int k = 0; k < 12; k++) v[k] = fma((double)pack_lo16(q, k), 1.0 / PACK_LO, (double)h[k] / PACK_HI);
  } else {
    const uint32_t w[8] = {r.a.x, r.a.y, r.a.z, r.a.w, r.b.x, r.b.y, r.b.z, r.b.w};
    const int h[12] = {pack_s21<0>(w), pack_s21<1>(w), pack_s21<2>(w),  pack_s21<3>(w),
                       pack_s21<4>(w), pack_s21<5>(w), pack_s21<6>(w),  pack_s21<7>(w),
                       pack_s21<8>(w), pack_s21<9>(w), pack_s21<10>(w), pack_s21<11>(w)};
#pragma unroll
    for (int k = 0; k < 12; k++) v[k] = NOSCALE ? (R)h[k] : (R)h[k] * R(1.0 / PACK21);
  }
#pragma unroll
  for (int c3 = 0; c3 < 3; c3++) {
    r0[c3] = cx<R>(v[2 * c3], v[2 * c3 + 1]);
    r1[c3] = cx<R>(v[6 + 2 * c3], v[6 + 2 * c3 + 1]);
  }
}
// Fixed-point operands of the FP64 polish's star taken as their integers q
// (2^31 times each real, exact in FP64) instead of q 2^-31: a matrix's rows
// 0, 1 then carry 2^31 and its rebuilt row 2 2^62.  raw_m_q brings row 2 to
// 2^31 (3 complex multiplies by 2^-31 instead of 12 real ones on the
// rows); raw_m_dagger_q gives columns 0, 1 at 2^31 and column 2 at 2^62, or
// at 2^31 with TO31.  U g^+ and U^+(x-mu) g^+(x-mu) formed from these carry
// 2^62 in columns 0, 1 and 2^93 in column 2 whatever the k of each product
// term, so the star is taken off that scale once at the end.  Powers of two
// scale every rounding exactly (the values stay far from overflow and
// underflow): the star is the same bit for bit as from the scaled reals.
template <typename R>
__device__ __forceinline__ void fx_rows_q(const Raw<const int4*, R>& r, Cx<R> r0[3], Cx<R> r1[3]) {
  const int4 a = r.v[0], b = r.v[1], c = r.v[2];
  r0[0] = cx<R>((R)a.x, (R)a.y);
  r0[1] = cx<R>((R)a.z, (R)a.w);
  r0[2] = cx<R>((R)b.x, (R)b.y);
  r1[0] = cx<R>((R)b.z, (R)b.w);
  r1[1] = cx<R>((R)c.x, (R)c.y);
  r1[2] = cx<R>((R)c.z, (R)c.w);
}
template <typename R>
__device__ __forceinline__ M3<R> raw_m_q(const Raw<const int4*, R>& r) {
  const R sc = R(1.0 / FX_SCALE);
  M3<R> m;
  fx_rows_q<R>(r, m.e[0], m.e[1]);
  m_row2(m);
#pragma unroll
  for (int c = 0; c < 3; c++) m.e[2][c] = cx<R>(m.e[2][c].re * sc, m.e[2][c].im * sc);
  return m;
}
template <typename R, bool TO31>
__device__ __forceinline__ M3<R> raw_m_dagger_q(const Raw<const int4*, R>& r) {
  const R sc = R(1.0 / FX_SCALE);
  Cx<R> r0[3], r1[3];
  fx_rows_q<R>(r, r0, r1);
  M3<R> d;
  m_dagger_from_rows(d, r0, r1);
  if constexpr (TO31) {
#pragma unroll
    for (int i = 0; i < 3; i++) d.e[i][2] = cx<R>(d.e[i][2].re * sc, d.e[i][2].im * sc);
  }
  return d;
}
// the star accumulated from raw_m_q / raw_m_dagger_q operands, back to the
// scale of the reals
template <typename R>
__device__ __forceinline__ void fx_star_unscale(M3<R>& K) {
  const R s2 = R(1.0 / (FX_SCALE * FX_SCALE)), s3 = R(1.0 / (FX_SCALE * FX_SCALE * FX_SCALE));
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      const R f = j < 2 ? s2 : s3;
      K.e[i][j] = cx<R>(K.e[i][j].re * f, K.e[i][j].im * f);
    }
}

template <typename R, typename LK>
__device__ __forceinline__ M3<R> raw_m(const Raw<LK, R>& r) {
  M3<R> m;
  raw_rows<R>(r, m.e[0], m.e[1]);
  m_row2(m);
  return m;
}
template <typename R, typename LK>
__device__ __forceinline__ M3<R> raw_m_dagger(const Raw<LK, R>& r) {
  Cx<R> r0[3], r1[3];
  raw_rows<R>(r, r0, r1);
  M3<R> d;
  m_dagger_from_rows(d, r0, r1);
  return d;
}

// rows 0, 1 of link field `field` at site index i, from plain planes or
// from packed planes (R = float: hi only; R = double: hi + lo)
template <typename R, bool NT, typename V>
__device__ __forceinline__ void link_rows(const V* U, int field, int stride, int i, Cx<R> r0[3],
                                          Cx<R> r1[3]) {
  Store<V>::template load<R, NT>(field_ptr(U, field, stride), stride, i, r0, r1);
}
template <typename R, bool NT, bool NOSCALE = false>
__device__ __forceinline__ void link_rows(const PackedLinks& U, int field, int stride, int i,
                                          Cx<R> r0[3], Cx<R> r1[3]) {
  Raw<PackedLinks, R> r;
  raw_load<R, NT>(U, field, stride, i, r);
  raw_rows<R, NOSCALE>(r, r0, r1);
}
template <typename R, bool NT, typename LK>
__device__ __forceinline__ M3<R> link_m(const LK& U, int field, int stride, int i) {
  M3<R> m;
  link_rows<R, NT>(U, field, stride, i, m.e[0], m.e[1]);
  m_row2(m);
  return m;
}
template <typename R, bool NT, typename LK>
__device__ __forceinline__ M3<R> link_m_dagger(const LK& U, int field, int stride, int i) {
  Cx<R> r0[3], r1[3];
  link_rows<R, NT>(U, field, stride, i, r0, r1);
  M3<R> d;
  m_dagger_from_rows(d, r0, r1);
  return d;
}

// One lattice of a batch: its geometry and its fields (FP64 double2 or FP32
// float4 storage, by the phase of the run).  Gribov copies share U.  g[0] and
// g[1] are the even and odd fields of the gauge transformation; galt is a
// third field that a check sweep writes the new even field into, and the
// check's finalize swaps g[0] and galt unless the lattice stops there.
struct BatchProblem {
  Geom geo;
  const void* U;
  PackedLinks pk;  // the packed form of U (mixed runs on packed links)
  void* g[2];
  void* galt;
};

// The two star terms of direction mu, U_mu(x) g^+(x+mu) and
// U_mu^+(x-mu) g^+(x-mu), added to K in the short form: each is a product of
// two SU(3) matrices, so rows 0, 1 are formed and row 2 is rebuilt as
// conj(row0 x row1) (the second as (g(x-mu) U_mu(x-mu))^+).  Equal to the
// full products for unitary operands; about 25% fewer complex multiplies.
// NC = 0: one direction, each operand decoded as it is loaded; NC = 1, 2:
// NC directions from mu, every operand loaded (Raw) before any term is formed.
// SC: the link rows come without their fixed-point scale (2^20 times the
// link, packed FP32 sweeps): the product's rows 0, 1 then carry 2^20 and its
// rebuilt row 2 2^40, and each is added to K with a multiply-add by the
// inverse power of two instead of an add; the backward link's rebuilt row 2
// is brought to 2^20 first.  Scaling by powers of two is exact in floating
// point, so K is the same bit for bit, and 12 + 6 multiplies per direction
// are saved.
template <typename R, bool SC>
__device__ __forceinline__ void add_term(Cx<R>& k, const Cx<R>& p, bool conj, R sc) {
  if constexpr (SC) {
    k = cx<R>(fma(p.re, sc, k.re), conj ? fma(p.im, -sc, k.im) : fma(p.im, sc, k.im));
  } else {
    k = cadd(k, conj ? cconj(p) : p);
  }
}
template <typename R, bool SC = false>
__device__ __forceinline__ void short_term_fwd(M3<R>& K, const Cx<R> a[2][3], Cx<R> b[3][3]) {
  const R s1 = R(1.0 / PACK21), s2 = R(1.0 / (PACK21 * PACK21));
  Cx<R> p[3][3];
  row2_of(b[2], b[0], b[1]);
  rows_mul_dagger(p, a, b);
  row2_of(p[2], p[0], p[1]);
#pragma unroll
  for (int r = 0; r < 3; r++)
#pragma unroll
    for (int c = 0; c < 3; c++) add_term<R, SC>(K.e[r][c], p[r][c], false, r < 2 ? s1 : s2);
}
template <typename R, bool SC = false>
__device__ __forceinline__ void short_term_bwd(M3<R>& K, const Cx<R> a[2][3], Cx<R> b[3][3]) {
  const R s1 = R(1.0 / PACK21), s2 = R(1.0 / (PACK21 * PACK21));
  Cx<R> p[3][3];
  row2_of(b[2], b[0], b[1]);
  if constexpr (SC) {
#pragma unroll
    for (int c = 0; c < 3; c++) b[2][c] = cx<R>(b[2][c].re * s1, b[2][c].im * s1);
  }
  rows_mul(p, a, b);
  row2_of(p[2], p[0], p[1]);
#pragma unroll
  for (int r = 0; r < 3; r++)
#pragma unroll
    for (int c = 0; c < 3; c++) add_term<R, SC>(K.e[r][c], p[c][r], true, c < 2 ? s1 : s2);
}
// The two terms of direction mu; ALL = true loads the direction's four
// operands before any is decoded.
template <typename V, typename R, bool NT, typename LK, bool ALL = false>
__device__ __forceinline__ void star_terms_short(M3<R>& K, LK U, const V* __restrict__ gop,
                                                 const Geom& geo, const Site& s, int i,
                                                 int parity, int mu) {
  const int S = geo.stride;
  // packed links in FP32: decoded without their scale (see short_term_fwd)
  constexpr bool SC = std::is_same<LK, PackedLinks>::value && std::is_same<R, float>::value;
  Cx<R> a[2][3], b[3][3];
  const int jf = nb_fwd(geo, s, i, mu);
  const int jb = nb_bwd(geo, s, i, mu);
  if constexpr (ALL) {
    Raw<LK, R> lf, lb;
    Raw<const V*, R> gf, gb;
    raw_load<R, NT>(U, parity * 4 + mu, S, i, lf);
    raw_load<R, false>(gop, S, jf, gf);
    raw_load<R, false>(gop, S, jb, gb);
    raw_load<R, NT>(U, (1 - parity) * 4 + mu, S, jb, lb);
    __builtin_amdgcn_sched_barrier(0);
    if constexpr (SC)
      raw_rows<R, true>(lf, a[0], a[1]);
    else
      raw_rows<R>(lf, a[0], a[1]);
    raw_rows<R>(gf, b[0], b[1]);
    short_term_fwd<R, SC>(K, a, b);
    raw_rows<R>(gb, a[0], a[1]);
    if constexpr (SC)
      raw_rows<R, true>(lb, b[0], b[1]);
    else
      raw_rows<R>(lb, b[0], b[1]);
    short_term_bwd<R, SC>(K, a, b);
    return;
  }
  if constexpr (SC)
    link_rows<R, NT, true>(U, parity * 4 + mu, S, i, a[0], a[1]);
  else
    link_rows<R, NT>(U, parity * 4 + mu, S, i, a[0], a[1]);
  Store<V>::template load<R, false>(gop, S, jf, b[0], b[1]);
  short_term_fwd<R, SC>(K, a, b);
  Store<V>::template load<R, false>(gop, S, jb, a[0], a[1]);
  if constexpr (SC)
    link_rows<R, NT, true>(U, (1 - parity) * 4 + mu, S, jb, b[0], b[1]);
  else
    link_rows<R, NT>(U, (1 - parity) * 4 + mu, S, jb, b[0], b[1]);
  short_term_bwd<R, SC>(K, a, b);
}

// The whole short star, software-pipelined: the operands of directions 0
// and 1 are loaded, then the terms of each direction mu are formed and only
// after them are the operands of direction mu + 2 loaded into the buffer
// direction mu freed.  An empty asm statement that takes K as read and
// written, and hands back the site index the next loads are addressed with,
// keeps the compiler from issuing those loads before the terms are done (the
// fully unrolled star, left to the compiler, issues every load first: 181
// VGPRs, 2 waves per SIMD).  The packed FP32 sweep takes 85 VGPRs this way
// (5 waves per SIMD) and is about 10% faster at 32^4.  The terms are formed
// in the same order as in the unrolled star, so the results are the same.
template <typename V, typename R, bool NT, typename LK, int NDIR>
__device__ __forceinline__ void star_short_pipelined(M3<R>& K, LK U, const V* __restrict__ gop,
                                                     const Geom& geo, const Site& s, int i,
                                                     int parity) {
  const int S = geo.stride;
  constexpr bool SC = std::is_same<LK, PackedLinks>::value && std::is_same<R, float>::value;
  Raw<LK, R> lf[2], lb[2];
  Raw<const V*, R> gf[2], gb[2];
  auto load_dir = [&](int mu, int k) {
    const int jf = nb_fwd(geo, s, i, mu);
    const int jb = nb_bwd(geo, s, i, mu);
    raw_load<R, NT>(U, parity * 4 + mu, S, i, lf[k]);
    raw_load<R, false>(gop, S, jf, gf[k]);
    raw_load<R, false>(gop, S, jb, gb[k]);
    raw_load<R, NT>(U, (1 - parity) * 4 + mu, S, jb, lb[k]);
  };
  auto terms = [&](int k) {
    Cx<R> a[2][3], b[3][3];
    if constexpr (SC)
      raw_rows<R, true>(lf[k], a[0], a[1]);
    else
      raw_rows<R>(lf[k], a[0], a[1]);
    raw_rows<R>(gf[k], b[0], b[1]);
    short_term_fwd<R, SC>(K, a, b);
    raw_rows<R>(gb[k], a[0], a[1]);
    if constexpr (SC)
      raw_rows<R, true>(lb[k], b[0], b[1]);
    else
      raw_rows<R>(lb[k], b[0], b[1]);
    short_term_bwd<R, SC>(K, a, b);
  };
  load_dir(0, 0);
  load_dir(1, 1);
#pragma unroll
  for (int mu = 0; mu < NDIR; mu++) {
    terms(mu & 1);
    if (mu + 2 < NDIR) {
      // the terms of direction mu are complete before the loads of
      // direction mu + 2 issue: the empty asm takes K as read and written
      // and hands back the site index those loads are addressed with
      int ii = i;
      asm volatile(""
                   : "+v"(ii), "+v"(K.e[0][0].re), "+v"(K.e[0][0].im), "+v"(K.e[0][1].re),
                     "+v"(K.e[0][1].im), "+v"(K.e[0][2].re), "+v"(K.e[0][2].im),
                     "+v"(K.e[1][0].re), "+v"(K.e[1][0].im), "+v"(K.e[1][1].re),
                     "+v"(K.e[1][1].im), "+v"(K.e[1][2].re), "+v"(K.e[1][2].im),
                     "+v"(K.e[2][0].re), "+v"(K.e[2][0].im), "+v"(K.e[2][1].re),
                     "+v"(K.e[2][1].im), "+v"(K.e[2][2].re), "+v"(K.e[2][2].im));
      const int z = ii - i;
      const int jf = nb_fwd(geo, s, i, mu + 2) + z;
      const int jb = nb_bwd(geo, s, i, mu + 2) + z;
      const int k = mu & 1;
      raw_load<R, NT>(U, parity * 4 + mu + 2, S, ii, lf[k]);
      raw_load<R, false>(gop, S, jf, gf[k]);
      raw_load<R, false>(gop, S, jb, gb[k]);
      raw_load<R, NT>(U, (1 - parity) * 4 + mu + 2, S, jb, lb[k]);
    }
  }
}

}  // namespace gfx
