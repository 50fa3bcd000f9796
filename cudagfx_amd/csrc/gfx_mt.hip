// gfx_mt.hip -- per-thread Mersenne Twisters (MT607, dcmt parameter sets)
// and the Box-Muller transform, with the reference's output layout and state
// semantics (MersenneTwister.cu:142-330):
//   * twister r writes its k-th number to out[r + k*4096];
//   * numbers are ((float)x + 1.0f) / 2^32 in (0, 1];
//   * after a call the 19-word state is stored as it lies in the circular
//     buffer and the next call with initialized=1 restarts at index 0;
//     params[r].seed <- state[0];
//   * initialized=0 reseeds from params[r].seed with
//     mt[j] = 1812433253*(mt[j-1] ^ (mt[j-1] >> 30)) + j.
// The generator follows the published MT recurrence (Matsumoto & Nishimura
// 1998; n = 19, m = 9, r = 1 for the 607-bit period, tempering shifts
// 12/7/15/18 with per-twister masks b, c).
#include <hip/hip_runtime.h>

#include <stdint.h>

namespace gfx {

constexpr int MT_COUNT = 4096;
constexpr int MT_N = 19;
constexpr int MT_M = 9;

__device__ __forceinline__ uint32_t mt_temper(uint32_t x, uint32_t b, uint32_t c) {
  x ^= (x >> 12);
  x ^= (x << 7) & b;
  x ^= (x << 15) & c;
  x ^= (x >> 18);
  return x;
}

// One step at compile-time buffer position K; mt[] stays in registers.
template <int K>
__device__ __forceinline__ uint32_t mt_step(uint32_t (&mt)[MT_N], uint32_t& m1, uint32_t a) {
  constexpr int I1 = (K + 1) % MT_N;
  constexpr int IM = (K + MT_M) % MT_N;
  const uint32_t m0 = m1;
  m1 = mt[I1];
  uint32_t x = (m0 & 0xFFFFFFFEu) | (m1 & 0x1u);
  x = mt[IM] ^ (x >> 1) ^ ((x & 1u) ? a : 0u);
  mt[K] = x;
  return x;
}

template <int K>
__device__ __forceinline__ void mt_emit(uint32_t (&mt)[MT_N], uint32_t& m1, uint32_t a, uint32_t b,
                                        uint32_t c, float* __restrict__ out, int r, int k) {
  const uint32_t x = mt_temper(mt_step<K>(mt, m1, a), b, c);
  out[r + (size_t)k * MT_COUNT] = ((float)x + 1.0f) / 4294967296.0f;
}

template <int K>
__device__ __forceinline__ void mt_round(uint32_t (&mt)[MT_N], uint32_t& m1, uint32_t a, uint32_t b,
                                         uint32_t c, float* __restrict__ out, int r, int k0, int n) {
  if constexpr (K < MT_N) {
    if (k0 + K < n) {
      mt_emit<K>(mt, m1, a, b, c, out, r, k0 + K);
      mt_round<K + 1>(mt, m1, a, b, c, out, r, k0, n);
    }
  }
}

// params: MT_COUNT x {matrix_a, mask_b, mask_c, seed}; state: MT_COUNT x 19
__global__ __launch_bounds__(256) void k_mt_generate(uint32_t* __restrict__ params,
                                                     uint32_t* __restrict__ state,
                                                     float* __restrict__ out, int nper,
                                                     int initialized) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= MT_COUNT) return;
  const uint32_t a = params[4 * r + 0], b = params[4 * r + 1], c = params[4 * r + 2];
  uint32_t mt[MT_N];
  if (!initialized) {
    mt[0] = params[4 * r + 3];
#pragma unroll
    for (int j = 1; j < MT_N; j++) mt[j] = 1812433253u * (mt[j - 1] ^ (mt[j - 1] >> 30)) + (uint32_t)j;
  } else {
#pragma unroll
    for (int j = 0; j < MT_N; j++) mt[j] = state[(size_t)r * MT_N + j];
  }
  uint32_t m1 = mt[0];
  // a full round of 19 outputs returns the circular index to 0
  int k0 = 0;
  for (; k0 + MT_N <= nper; k0 += MT_N) mt_round<0>(mt, m1, a, b, c, out, r, k0, nper);
  if (k0 < nper) mt_round<0>(mt, m1, a, b, c, out, r, k0, nper);
  params[4 * r + 3] = mt[0];
#pragma unroll
  for (int j = 0; j < MT_N; j++) state[(size_t)r * MT_N + j] = mt[j];
}

// Cartesian Box-Muller on pairs (out[r + 2k*4096], out[r + (2k+1)*4096])
// (MersenneTwister.cu:209-229), one thread per pair, with the library's
// single-precision log, sqrt and sincos.
__global__ __launch_bounds__(256) void k_box_muller(float* __restrict__ d, int nper) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long npair = (long)MT_COUNT * (nper / 2);
  if (t >= npair) return;
  const int r = (int)(t % MT_COUNT);
  const long k = t / MT_COUNT;
  float* p1 = d + r + (2 * k) * MT_COUNT;
  float* p2 = d + r + (2 * k + 1) * MT_COUNT;
  const float u1 = *p1, u2 = *p2;
  const float rad = sqrtf(-2.0f * logf(u1));
  const float phi = 2 * 3.14159265358979f * u2;
  float sn, cs;
  sincosf(phi, &sn, &cs);
  *p1 = rad * cs;
  *p2 = rad * sn;
}

void launch_mt_generate(uint32_t* params, uint32_t* state, float* out, int nper, int initialized,
                        hipStream_t st) {
  hipLaunchKernelGGL(k_mt_generate, dim3(MT_COUNT / 256), dim3(256), 0, st, params, state, out,
                     nper, initialized);
}

void launch_box_muller(float* d, int nper, hipStream_t st) {
  const long npair = (long)MT_COUNT * (nper / 2);
  if (npair <= 0) return;
  hipLaunchKernelGGL(k_box_muller, dim3((unsigned)((npair + 255) / 256)), dim3(256), 0, st, d,
                     nper);
}

}  // namespace gfx
