// gfx_noalloc.hip -- gfx_calc_functional_dev: F, dAdA and max dAdA of
// caller-owned device fields without any allocation, for callers that cannot
// allocate device memory (code under HIP graph capture, serving loops,
// harnesses that forbid hipMalloc).
//
// The fields are in the reference's device layout (2-row double2 storage,
// su3manip.c:212-230, 281-299): link U_mu(s) at gf[6*(4 s + mu) + c] and
// g(s) at trafo[6 s + c], c = 0..5 the elements (0,0) (0,1) (0,2) (1,0)
// (1,1) (1,2), row 2 rebuilt as conj(row0 x row1); s lexicographic as
// cudagaugefix.cu:79-85.  This is the form in which the reference's
// calc_functional (overrelax.cu:610-671) took its device fields.
//
// The block partials of the reduction live in `scratch`, a file-scope device
// array of fixed size: the entry point needs no workspace argument and calls
// no allocating routine.  Each GPU holds its own copy of the array until the
// process ends; its address on the current device is looked up once per
// device with hipGetSymbolAddress and used for every launch (the host-side
// name of the array is not a device address).  Every block writes its own
// partials before the finalize reads them, so nothing has to be cleared
// between calls; calls that overlap on one GPU share the array and race.
#include <hip/hip_runtime.h>

#include <cstdio>

#include "../../include/cudagfx.h"
#include "gfx_lattice.h"
#include "gfx_su3.h"

namespace gfx {

constexpr int NA_BLK = 256;
// at most this many blocks of NA_BLK sites: 16,777,216 sites (64^4)
constexpr int NA_MAX_BLOCKS = 65536;
constexpr long NA_MAX_SITES = (long)NA_MAX_BLOCKS * NA_BLK;
constexpr int NA_MAX_DEVICES = 64;

// 3 doubles (sum F, sum dAdA, max dAdA) per block: 1.5 MiB
static __device__ __align__(256) double scratch[3 * NA_MAX_BLOCKS];

static void* scratch_ptr[NA_MAX_DEVICES] = {nullptr};

struct LexGeom {
  int L[4];
  int V;
  FastDiv d0, d1, d2;
};

__device__ __forceinline__ M3<double> na_load(const double2* __restrict__ f, long m) {
  M3<double> u;
  const double2* p = f + 6 * m;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const double2 a = p[c], b = p[3 + c];
    u.e[0][c] = cx<double>(a.x, a.y);
    u.e[1][c] = cx<double>(b.x, b.y);
  }
  m_row2(u);
  return u;
}

__device__ __forceinline__ M3<double> na_dagger(const M3<double>& u) {
  M3<double> d;
#pragma unroll
  for (int r = 0; r < 3; r++)
#pragma unroll
    for (int c = 0; c < 3; c++) d.e[r][c] = cconj(u.e[c][r]);
  return d;
}

__device__ __forceinline__ double na_wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ __forceinline__ double na_wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}

// block sums of a, b and max of c, valid in thread 0
__device__ __forceinline__ void na_block_reduce(double& a, double& b, double& c) {
  __shared__ double red[3][NA_BLK / 64];
  a = na_wave_sum(a);
  b = na_wave_sum(b);
  c = na_wave_max(c);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[0][w] = a;
    red[1][w] = b;
    red[2][w] = c;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < NA_BLK / 64; k++) {
      a += red[0][k];
      b += red[1][k];
      c = fmax(c, red[2][k]);
    }
  }
}

// per site: F(s) = sum_mu Re tr g(s) U_mu(s) g^+(s+mu) and Re tr Delta Delta^+
// with Delta(s) = sum_mu A(g(s-mu) U_mu(s-mu) g^+(s)) - A(g(s) U_mu(s)
// g^+(s+mu)), in the order of dev_functional (overrelax.cu:482-605)
__global__ __launch_bounds__(NA_BLK) void k_functional_2v(const double2* __restrict__ gf,
                                                          const double2* __restrict__ trafo,
                                                          LexGeom geo, double* __restrict__ part) {
  const int s = blockIdx.x * NA_BLK + threadIdx.x;
  double F = 0.0, d = 0.0;
  if (s < geo.V) {
    int x[4];
    const uint32_t q0 = fdiv((uint32_t)s, geo.d0);
    x[0] = s - (int)q0 * geo.L[0];
    const uint32_t q1 = fdiv(q0, geo.d1);
    x[1] = (int)(q0 - q1 * (uint32_t)geo.L[1]);
    const uint32_t q2 = fdiv(q1, geo.d2);
    x[2] = (int)(q1 - q2 * (uint32_t)geo.L[2]);
    x[3] = (int)q2;
    const int st[4] = {1, geo.L[0], geo.L[0] * geo.L[1], geo.L[0] * geo.L[1] * geo.L[2]};
    M3<double> gather;
    m_zero(gather);
#pragma unroll 1
    for (int mu = 0; mu < 4; mu++) {
      const int sf = (x[mu] == geo.L[mu] - 1) ? s - (geo.L[mu] - 1) * st[mu] : s + st[mu];
      const int sb = (x[mu] == 0) ? s + (geo.L[mu] - 1) * st[mu] : s - st[mu];
      const M3<double> G = na_load(trafo, s);
      M3<double> help = m_mul(G, na_load(gf, 4L * s + mu));
      M3<double> fw = m_mul(help, na_dagger(na_load(trafo, sf)));
      F += m_retrace(fw);
      help = m_mul(na_load(trafo, sb), na_load(gf, 4L * sb + mu));
      M3<double> bw = m_mul(help, na_dagger(G));
      vecpot(fw);
      vecpot(bw);
#pragma unroll
      for (int a = 0; a < 3; a++)
#pragma unroll
        for (int c = 0; c < 3; c++) gather.e[a][c] = cadd(gather.e[a][c], csub(bw.e[a][c], fw.e[a][c]));
    }
    d = m_retrace(m_mul_d(gather, gather));
  }
  double mx = d;
  na_block_reduce(F, d, mx);
  if (threadIdx.x == 0) {
    double* o = part + 3 * (size_t)blockIdx.x;
    o[0] = F;
    o[1] = d;
    o[2] = mx;
  }
}

// grid pass: result = F / (4 V), dAdA / V, max dAdA (overrelax.cu:647-661)
__global__ __launch_bounds__(NA_BLK) void k_functional_2v_finalize(const double* __restrict__ part,
                                                                   int nb, double V,
                                                                   double* __restrict__ result) {
  double F = 0.0, d = 0.0, mx = 0.0;
  for (int b = threadIdx.x; b < nb; b += NA_BLK) {
    F += part[3 * (size_t)b];
    d += part[3 * (size_t)b + 1];
    mx = fmax(mx, part[3 * (size_t)b + 2]);
  }
  na_block_reduce(F, d, mx);
  if (threadIdx.x == 0) {
    result[0] = F / (4.0 * V);
    result[1] = d / V;
    result[2] = mx;
  }
}

// the device address of `scratch` on the current device, looked up once per
// device
static int scratch_on_device(double** out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NA_MAX_DEVICES) return GFX_ERR_DEVICE;
  if (!scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      fprintf(stderr, "cudagfx: hipGetSymbolAddress(scratch): %s\n", hipGetErrorString(e));
      return GFX_ERR_DEVICE;
    }
    scratch_ptr[dev] = p;
  }
  *out = static_cast<double*>(scratch_ptr[dev]);
  return GFX_OK;
}

}  // namespace gfx

using namespace gfx;

extern "C" int gfx_calc_functional_dev(const double* gf, const double* trafo, int lx, int ly,
                                       int lz, int lt, double* result, void* stream) {
  if (!gf || !trafo || !result || lx < 1 || ly < 1 || lz < 1 || lt < 1) return GFX_ERR_GEOMETRY;
  const long V = (long)lx * ly * lz * lt;
  if (V > NA_MAX_SITES) {
    fprintf(stderr,
            "cudagfx: gfx_calc_functional_dev: %ld sites exceed the %ld of its static scratch\n", V,
            NA_MAX_SITES);
    return GFX_ERR_SIZE;
  }
  double* part = nullptr;
  int rc = scratch_on_device(&part);
  if (rc) return rc;
  LexGeom geo;
  geo.L[0] = lx;
  geo.L[1] = ly;
  geo.L[2] = lz;
  geo.L[3] = lt;
  geo.V = (int)V;
  geo.d0 = make_fastdiv(lx);
  geo.d1 = make_fastdiv(ly);
  geo.d2 = make_fastdiv(lz);
  const int nb = (int)((V + NA_BLK - 1) / NA_BLK);
  hipStream_t st = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_functional_2v, dim3(nb), dim3(NA_BLK), 0, st,
                     reinterpret_cast<const double2*>(gf), reinterpret_cast<const double2*>(trafo),
                     geo, part);
  hipLaunchKernelGGL(k_functional_2v_finalize, dim3(1), dim3(NA_BLK), 0, st, part, nb, (double)V,
                     result);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    fprintf(stderr, "cudagfx: gfx_calc_functional_dev: %s\n", hipGetErrorString(e));
    return GFX_ERR_DEVICE;
  }
  return GFX_OK;
}
