// gfx_internal.h -- launchers shared between the kernel and API translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "gfx_lattice.h"

namespace gfx {

// sweep variants: bits 0-2 pick the kernel (see k_or_sweep: block size,
// rolled/unrolled star loop, non-temporal link loads), bits 3-4 the
// workgroup order (block_order tiling 0..3)
constexpr int KERNEL_VARIANTS = 8;
constexpr int TILINGS = 4;
constexpr int SWEEP_VARIANTS = KERNEL_VARIANTS * TILINGS;
int sweep_variant_block(int v);
bool sweep_variant_rolled(int v);
bool sweep_variant_nt(int v);
inline int sweep_variant_tiling(int v) { return v >> 3; }
inline int block_index(int block) { return block == 256 ? 0 : 1; }

// Workgroup order of a sweep over one lattice (local block indices, block
// sites each).  tile 0: lexicographic.  tiles 1, 2: a superrow is the G
// consecutive (x2, x3) planes (h0*L1 sites each) that make a whole number of
// blocks; superrows are taken in chunks of 4 (tile 1) or 8 (tile 2)
// consecutive x2-superrows, x3 outermost inside a chunk, so that the x3
// neighbours of a block were read a few superrows earlier and are still in
// L2.  tile 3: the planes are also cut in x1, into pieces of about 512 sites,
// for large lattices whose planes are too big for that.  Returns false (and
// the lexicographic order) when the order does not apply to the lattice.
bool block_order(const Geom& geo, int block, int tile, std::vector<int>& order);

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own
// L2): launch position p runs on XCD p % 8.  Deal `in` so that XCD k works
// through the k-th contiguous eighth of it in order.
template <typename T>
void deal_to_xcds(const std::vector<T>& in, std::vector<T>& out) {
  const size_t n = in.size();
  out.resize(n);
  size_t start = 0;
  for (size_t k = 0; k < 8; k++) {
    const size_t cnt = n / 8 + (k < n % 8 ? 1 : 0);
    for (size_t j = 0; j < cnt; j++) out[k + 8 * j] = in[start + j];
    start += cnt;
  }
}
// one half sweep over the fields ge (even) and go (odd) of g.  chk (Landau
// only): 0 plain sweep; 1 check after the update (odd sweep of a check
// iteration); 2 check before the update (even sweep that follows it).  The
// new field of the updated parity goes to galt when given (always with chk
// = 2), otherwise the sweep is in place.  Checks write 3 doubles per block
// to partials (sweep_blocks(nh, variant) blocks).
// bmap: launch position -> local block (block_order dealt to the XCDs);
// nblocks > 0 launches only the first nblocks entries (a map of a subset)
void launch_sweep_f64(const double2* U, double2* ge, double2* go, const Geom& geo, int parity,
                      int ndir, double omega, const unsigned char* active, int variant,
                      const int* bmap, hipStream_t st, int chk = 0, double2* galt = nullptr,
                      double* partials = nullptr, int nblocks = 0);
void launch_sweep_f32(const float4* U, float4* ge, float4* go, const Geom& geo, int parity,
                      int ndir, double omega, const unsigned char* active, int variant,
                      const int* bmap, hipStream_t st, int chk = 0, float4* galt = nullptr,
                      double* partials = nullptr, int nblocks = 0);
// half sweeps on packed links (Landau and Coulomb, arguments as above): FP32
// g on the 21-bit links (every variant), FP64 g on 24-bit high + 16-bit low
// parts (rolled variants only, bit 1 of the variant ignored)
void launch_sweep_packed_f32(const PackedArgs& p, float4* ge, float4* go, const Geom& geo,
                             int parity, int ndir, double omega, const unsigned char* active,
                             int variant, const int* bmap, hipStream_t st, int chk = 0,
                             float4* galt = nullptr, double* partials = nullptr, int nblocks = 0);
void launch_sweep_packed_f64(const PackedArgs& p, double2* ge, double2* go, const Geom& geo,
                             int parity, int ndir, double omega, const unsigned char* active,
                             int variant, const int* bmap, hipStream_t st, int chk = 0,
                             double2* galt = nullptr, double* partials = nullptr, int nblocks = 0);
// the FP64 polish of mixed Landau and Coulomb runs on 32-bit fixed-point
// links and g (int4 planes, Store<int4>), arguments as launch_sweep_f64
void launch_sweep_fx(const int4* U, int4* ge, int4* go, const Geom& geo, int parity, int ndir,
                     double omega, const unsigned char* active, int variant, const int* bmap,
                     hipStream_t st, int chk = 0, int4* galt = nullptr, double* partials = nullptr,
                     int nblocks = 0);
int sweep_blocks(int nh, int variant);
// per-slice grid pass of a fused Coulomb check (bps blocks per slice; only
// active slices are written)
void launch_slice_finalize(const double* podd, const double* peven, int bps, int nt,
                           const unsigned char* active, double normF, double normD, double eps,
                           double* out, hipStream_t st);
// grid pass of a fused check: out[0..3] = F/normF, dAdA/normD, max, max < eps
void launch_check_finalize(const double* podd, const double* peven, int nb, double normF,
                           double normD, double eps, double* out, hipStream_t st);
int functional_blocks(int nsh);
int functional_block_size();
void launch_stream_copy(const void* src, void* dst, size_t bytes, hipStream_t st);

// batched lattices (gfx_batch_api.hip)
// the FP64 problems' fields: FP64 links and g (BLINK_F64), packed links and
// FP64 g (BLINK_PACKED), or 32-bit fixed-point links and g (BLINK_FIXED);
// the FP32 problems read FP32 links (BLINK_F64) or the packed 21-bit links
enum { BLINK_F64 = 0, BLINK_PACKED = 1, BLINK_FIXED = 2 };
void launch_batch_sweep(bool fp32, const BatchProblem* probs, const int2* bmap, int nblocks,
                        const int* active, const int* poff, int parity, double omega, int variant,
                        int chk, double* partials, hipStream_t st, int links);
// device arrays of a batch check, index 0 = FP64 phase, 1 = FP32 phase
struct BatchCheck {
  const double* pod[2];  // block partials of the odd check sweeps
  const double* pev[2];  // and of the even ones
  const int* poff[2];    // partial offsets per lattice (block size of the phase's variant)
  const int* inmap[2];   // per lattice: its blocks are in the precision's current map
  BatchProblem* probs[2];
  int* act[2];
  int* conv;
  int* sw;
  double* results;
  int* iter;
  double* best;
  int* stall;
  int links;  // BLINK_*: the FP64 stops take the packing or fixed-point bound
};
// the persistent form of the check periods (gfx_persist.hip)
struct PersistArgs {
  const int2* bmap;  // the workgroup map of the precision swept
  int nblocks;
  double* pod;  // check partials of the odd sweeps of that precision
  double* pev;  // and of the even ones
  int nprob;
  int ck;       // iterations per check period
  int periods;  // check periods at most
  double eps, switch_theta;
  int maxit, stall_checks;
  double stall_gate;
  int* changed;  // set when a lattice stops or leaves FP32 (cleared by the host)
  int* done;     // check periods run
};
hipError_t launch_batch_persistent(bool fp32, int links, int variant, const PersistArgs& a,
                                   const BatchCheck& c, double omega, hipStream_t st);
bool persistent_supported();
void launch_batch_finalize(const BatchCheck& c, int nprob, int ndir, double eps,
                           double switch_theta, int maxit, int stall_checks, double stall_gate,
                           hipStream_t st);
void launch_batch_convert(const BatchProblem* probs32, const BatchProblem* probs64,
                          const int2* bmap, int nblocks, const int* conv, hipStream_t st,
                          bool fixed = false);
int batch_convert_block();
void launch_batch_advance(int* iter, int step, hipStream_t st);
// batched annealing (gfx_heatbath.hip); bmap of sa_block_size() blocks
// fp32: the FP32 problems (FP32 links and g, FP32 arithmetic; sampling in FP64)
void launch_batch_sa_sweep(const BatchProblem* probs, const int2* bmap, int nblocks,
                           const int* roff, int parity, const double* betas, const int* step,
                           int sampler, const float* gauss, const float* unif, hipStream_t st,
                           bool fp32 = false);
void launch_batch_micro_sweep(const BatchProblem* probs, const int2* bmap, int nblocks, int parity,
                              hipStream_t st, bool fp32 = false);
int sa_block_size();
int kp_gauss_per_site();
int kp_unif_per_site();
void launch_fill_i32(int* p, int n, int v, hipStream_t st);
void launch_fill_f64(double* p, int n, double v, hipStream_t st);
void launch_functional_f64(const double2* U, const double2* ge, const double2* go, const Geom& geo,
                           int ndir, int nseg, double* partials, double eps, double* out,
                           hipStream_t st);
void launch_functional_f32(const float4* U, const float4* ge, const float4* go, const Geom& geo,
                           int ndir, int nseg, double* partials, double eps, double* out,
                           hipStream_t st);
void launch_plaquette_f64(const double2* U, const Geom& geo, double* partials, double* out,
                          hipStream_t st);
void launch_pack_f64(const double* aos, double2* dst, const Geom& geo, int s0, int ns, int nmat,
                     hipStream_t st);
void launch_unpack_f64(const double2* src, double* aos, const Geom& geo, int s0, int ns, int nmat,
                       hipStream_t st);
void launch_apply_trafo_aos(double* aos, const double2* g, const Geom& geo, int s0, int ns,
                            hipStream_t st);
void launch_f64_to_f32(const double2* src, float4* dst, const Geom& geo, int nfield,
                       hipStream_t st);
// packed form of nfield FP64 link fields
void launch_pack_links(const double2* src, const PackedArgs& p, const Geom& geo, int nfield,
                       hipStream_t st);
void launch_g32_to_g64(const float4* src, double2* dst, const Geom& geo, hipStream_t st);
// fixed-point fields: nfield FP64 fields -> fixed point; the two g fields
// (even, odd) fixed point -> FP64 and FP32 -> fixed point, reunitarised in FP64
void launch_f64_to_fx(const double2* src, int4* dst, const Geom& geo, int nfield, hipStream_t st);
void launch_fx_to_g64(const int4* src, double2* dst, const Geom& geo, hipStream_t st);
void launch_g32_to_fx(const float4* src, int4* dst, const Geom& geo, hipStream_t st);
// sites [i0, i0 + n) of both parities, FP32 -> fixed point, fields given separately
void launch_g32_to_fx_range(const float4* se, const float4* so, int4* de, int4* dd,
                            const Geom& geo, int i0, int n, hipStream_t st);
// sites [i0, i0 + n) of both parities, fields given separately
void launch_g32_to_g64_range(const float4* se, const float4* so, double2* de, double2* dd,
                             const Geom& geo, int i0, int n, hipStream_t st);

}  // namespace gfx
