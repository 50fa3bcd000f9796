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

// A call continues the linear sequence y with y_0..y_18 = the stored buffer
// (position K holds y_K; "restart at index 0") and
//   y_{t+19} = y_{t+9} ^ A(y_t & upper | y_{t+1} & lower),
// emitting temper(y_{19+k}) for k < nper; afterwards position K of the
// buffer holds the newest y_t with t = K (mod 19).  y_{19+k} .. y_{28+k}
// depend only on y_k .. y_{k+18}, so MT_L = 10 lanes per twister produce ten
// outputs per step from a 32-entry ring in LDS (one barrier per step; the
// slots written in a step are never read in it).  A block holds MT_TPB
// twisters; lane (j, r) = thread j * MT_TPB + r, so a wavefront writes
// MT_TPB consecutive floats of four output rows.  16 twisters per block
// spread the 4096 twisters over all 256 CUs (16% faster than 32 per block;
// 8 per block, or the ten lanes of a twister inside one wavefront with no
// block barrier, are not faster: a step is bound by the latency of its LDS
// round trip, not by the barrier).
constexpr int MT_TPB = 16;
constexpr int MT_L = 10;

// params: MT_COUNT x {matrix_a, mask_b, mask_c, seed}; state: MT_COUNT x 19
__global__ __launch_bounds__(MT_TPB * MT_L) void k_mt_generate(uint32_t* __restrict__ params,
                                                               uint32_t* __restrict__ state,
                                                               float* __restrict__ out, int nper,
                                                               int initialized) {
  __shared__ uint32_t ring[MT_TPB][33];  // 33: the twisters of a wavefront hit different banks
  const int rl = threadIdx.x % MT_TPB, j = threadIdx.x / MT_TPB;
  const int r = blockIdx.x * MT_TPB + rl;
  const uint32_t a = params[4 * r + 0], b = params[4 * r + 1], c = params[4 * r + 2];
  if (!initialized) {
    if (j == 0) {
      uint32_t y = params[4 * r + 3];
      ring[rl][0] = y;
      for (int t = 1; t < MT_N; t++) {
        y = 1812433253u * (y ^ (y >> 30)) + (uint32_t)t;
        ring[rl][t] = y;
      }
    }
  } else {
    for (int t = j; t < MT_N; t += MT_L) ring[rl][t] = state[(size_t)r * MT_N + t];
  }
  __syncthreads();
  for (int k0 = 0; k0 < nper; k0 += MT_L) {
    const int k = k0 + j;
    if (k < nper) {
      const uint32_t y0 = ring[rl][k & 31], y1 = ring[rl][(k + 1) & 31];
      uint32_t x = (y0 & 0xFFFFFFFEu) | (y1 & 0x1u);
      x = ring[rl][(k + MT_M) & 31] ^ (x >> 1) ^ ((x & 1u) ? a : 0u);
      ring[rl][(k + MT_N) & 31] = x;
      out[r + (size_t)k * MT_COUNT] = ((float)mt_temper(x, b, c) + 1.0f) / 4294967296.0f;
    }
    __syncthreads();
  }
  const int T = MT_N - 1 + nper;  // index of the last y produced
  for (int K = j; K < MT_N; K += MT_L) {
    const uint32_t v = ring[rl][(T - (T - K) % MT_N) & 31];
    state[(size_t)r * MT_N + K] = v;
    if (K == 0) params[4 * r + 3] = v;
  }
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
  hipLaunchKernelGGL(k_mt_generate, dim3(MT_COUNT / MT_TPB), dim3(MT_TPB * MT_L), 0, st, params,
                     state, out, nper, initialized);
}

void launch_box_muller(float* d, int nper, hipStream_t st) {
  const long npair = (long)MT_COUNT * (nper / 2);
  if (npair <= 0) return;
  hipLaunchKernelGGL(k_box_muller, dim3((unsigned)((npair + 255) / 256)), dim3(256), 0, st, d,
                     nper);
}

}  // namespace gfx
