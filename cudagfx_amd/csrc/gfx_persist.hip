// gfx_persist.hip -- the persistent form of the batching form's check
// periods: one cooperative launch that keeps the lattices of a batch
// sweeping in one precision, check period after check period, until a
// lattice stops or leaves FP32, all have stopped, or the launch's periods
// are done.  The workgroups stay resident and take the (lattice, block)
// items of each half sweep from the precision's workgroup map with a grid
// stride; a grid barrier separates the half sweeps (the red-black order),
// the check sweep from the per-lattice check, and that check from the next
// period.  The same work as one replay of the check-period graph
// (gfx_batch_api.hip), with the per-lattice check inside the launch.
#include <hip/hip_cooperative_groups.h>
#include <hip/hip_runtime.h>

#include <algorithm>

#include "gfx_internal.h"
#include "gfx_sweep.h"

namespace gfx {

namespace cg = cooperative_groups;

// The grid is a whole number of rounds of the 8 XCDs, so item w of a map
// runs on XCD w % 8 as it would as workgroup w of a plain launch: the map's
// dealing of blocks to XCDs holds.  The three forms of the sweep item
// (plain, odd check, even check) share one call site with the form chosen
// at run time: with a call site per form the compiler allocated 223-226
// VGPRs for the fixed-point FP64 form (2 waves per SIMD) where its plain
// launches take 168-170.
template <typename V, typename R, int BLOCK, bool ROLLED, bool PK>
__global__ __launch_bounds__(BLOCK) void k_batch_persistent(PersistArgs a, BatchCheck c, R c1,
                                                            R c2) {
  constexpr int pr = std::is_same<R, float>::value ? 1 : 0;
  cg::grid_group grid = cg::this_grid();
  const BatchProblem* probs = c.probs[pr];
  const int* act = c.act[pr];
  const int* poff = c.poff[pr];
  for (int per = 0; per < a.periods; per++) {
    // the 2 ck half sweeps of a period, one body for all (one call site:
    // registers as in the plain launches): h = 0 the odd half of its first
    // iteration (the even half ran as the check sweep of the period
    // before), then even and odd halves, the odd half of the last iteration
    // carrying the check (h = 2 ck - 2), and the even half of the next
    // iteration carrying the check (h = 2 ck - 1)
    for (int h = 0; h < 2 * a.ck; h++) {
      const int parity = (h & 1) ? 0 : 1;
      const int chk = h == 2 * a.ck - 2 ? 1 : h == 2 * a.ck - 1 ? 2 : 0;
      double* partials = chk == 2 ? a.pev : a.pod;
      for (int w = blockIdx.x; w < a.nblocks; w += gridDim.x) {
        batch_sweep_item<V, R, 4, BLOCK, ROLLED, false, 3, PK>(probs, a.bmap[w], act, poff, parity,
                                                               c1, c2, partials, chk);
        // the block reduction's LDS is used again by the next item
        if (chk) __syncthreads();
      }
      grid.sync();
    }
    for (int q = blockIdx.x; q < a.nprob; q += gridDim.x) {
      batch_finalize_lattice(q, c, 4, a.eps, a.switch_theta, a.maxit, a.stall_checks,
                             a.stall_gate, a.changed);
      __syncthreads();
    }
    grid.sync();
    // every workgroup reads the flag after the barrier that follows its last
    // write: all leave together or none does
    const int changed = __hip_atomic_load(a.changed, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      __hip_atomic_fetch_add(c.iter, a.ck, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      *a.done = per + 1;
    }
    if (changed) break;
  }
}

template <typename V, typename R, int BLOCK, bool ROLLED, bool PK>
static hipError_t launch_persistent_v(const PersistArgs& a, const BatchCheck& c, double omega,
                                      hipStream_t st) {
  const void* kern = reinterpret_cast<const void*>(&k_batch_persistent<V, R, BLOCK, ROLLED, PK>);
  int per_cu = 0, dev = 0, ncu = 0;
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, BLOCK, 0);
  if (e != hipSuccess) return e;
  if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
  if ((e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess)
    return e;
  // resident workgroups, a whole number of XCD rounds, no more than the map
  // needs (rounded up to the rounds)
  int grid = per_cu * ncu / 8 * 8;
  grid = std::min(grid, (a.nblocks + 7) / 8 * 8);
  if (grid < 8) return hipErrorInvalidConfiguration;
  PersistArgs aa = a;
  BatchCheck cc = c;
  R c1 = (R)((1.0 + omega - 1) / 1);
  R c2 = (R)((1.0 + omega - 2) / 2);
  void* args[] = {&aa, &cc, &c1, &c2};
  return hipLaunchCooperativeKernel(kern, dim3(grid), dim3(BLOCK), args, 0, st);
}

template <typename V, typename R, bool PK, bool ROLLED_ONLY>
static hipError_t launch_persistent_t(const PersistArgs& a, const BatchCheck& c, double omega,
                                      int variant, hipStream_t st) {
  const bool b64 = sweep_variant_block(variant) == 64;
  if (ROLLED_ONLY || sweep_variant_rolled(variant))
    return b64 ? launch_persistent_v<V, R, 64, true, PK>(a, c, omega, st)
               : launch_persistent_v<V, R, 256, true, PK>(a, c, omega, st);
  if constexpr (!ROLLED_ONLY)
    return b64 ? launch_persistent_v<V, R, 64, false, PK>(a, c, omega, st)
               : launch_persistent_v<V, R, 256, false, PK>(a, c, omega, st);
  return hipErrorInvalidValue;
}

hipError_t launch_batch_persistent(bool fp32, int links, int variant, const PersistArgs& a,
                                   const BatchCheck& c, double omega, hipStream_t st) {
  if (fp32)
    return links != BLINK_F64 ? launch_persistent_t<float4, float, true, false>(a, c, omega, variant, st)
                              : launch_persistent_t<float4, float, false, false>(a, c, omega, variant, st);
  if (links == BLINK_FIXED)  // fixed-point fields: rolled only, as their plain launches
    return launch_persistent_t<int4, double, false, true>(a, c, omega, variant, st);
  if (links == BLINK_PACKED)
    return launch_persistent_t<double2, double, true, false>(a, c, omega, variant, st);
  return launch_persistent_t<double2, double, false, false>(a, c, omega, variant, st);
}

bool persistent_supported() {
  int dev = 0, coop = 0;
  if (hipGetDevice(&dev) != hipSuccess) return false;
  if (hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev) != hipSuccess) return false;
  return coop != 0;
}

}  // namespace gfx
