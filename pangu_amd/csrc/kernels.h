// Launch interface between the C-ABI layer and the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "common.h"
#include "physics_dev.h"

namespace pangu {

// Per-launch constants (passed by value; lives in kernel arguments).
struct KCtx {
  const DevBlock *blk;
  const DevGeo *geo;
  int nblocks;
  int np, nc, iBX, iENT, iKEL;
  int ndim;
  int nk;   // interior (or launch) planes per block for blockIdx.z decoding
  int nkg;  // ghosted planes per block
  Phys ph;
};

// Device pointer tables of one container: prim/cons [nblocks*ncomp],
// flux[d] [nblocks*ncons].
struct KFields {
  double *const *prim;
  double *const *cons;
  double *const *flux[3];
};

struct SharedTabs {
  double *const *flag;
  double *const *flag_rec;
  int32_t *const *status;
  int32_t *const *iters;
};

// Launch range of a kernel over every block.  order: 0 = plane-major tiles
// (consecutive tiles sweep a k plane), 1 = column-major tiles (consecutive
// tiles walk k for one 256-cell run of the (j, i) plane, so the x3-independent
// geometry of that run is reused from L2 by every plane).
struct Range3 {
  int lo[3], hi[3], cnt[3];
  int order;
};

// Optional per-launch clock: when a LaunchCtx carries one, every kernel
// launch is followed by a HIP event on the launch stream, so the duration of
// launch i is the time between events i and i+1.  Timing runs only: not for
// code under graph capture.
struct KernelClock {
  std::vector<const char *> names;  // names[i]: the launch ended by events[i + 1]
  std::vector<hipEvent_t> events;
  size_t n = 0;                     // events recorded since reset()
  hipError_t mark(const char *name, hipStream_t s);
  void reset() { n = 0; }
  ~KernelClock();
};

// Host-side launch context: kernel constants plus the (uniform) block shape.
struct LaunchCtx {
  KCtx k;
  DevBlock hblk;
  int limiter, riemann, metric, recovery;
  int bc[3][2];
  int nb[3];
  int order[PANGU_KF_COUNT] = {};  // tile order per kernel family (PANGU_ORDER_*)
  int split_source = 0;            // fused stage: source term in a kernel of its own
  KernelClock *clock = nullptr;
};

// Attaches (or, with nullptr, detaches) a clock to a context's launches.
void set_kernel_clock(pangu_ctx *c, KernelClock *clock);

int launch_fix_cons(const LaunchCtx &L, const KFields &F, uint8_t *const *mask, bool fix, bool cons,
                    hipStream_t s);
int launch_fluxes(const LaunchCtx &L, const KFields &F, double *const *alfven, hipStream_t s);
int launch_ct(const LaunchCtx &L, const KFields &F, double *const *emf, hipStream_t s);
// mode: 1 divergence -> dudt, 2 update from dudt, 3 source only, 4 fused
// divergence + update + source, 5 fused divergence + update without source
int launch_div_update(const LaunchCtx &L, int mode, const KFields &F0, const KFields &Finit,
                      const KFields &F1, double *const *dudt, const double *dt_dev, double beta,
                      double bdt_host, hipStream_t s);
// rec: 0 chainer, 1 kastaun, 2 1dvsq, -1 none (heating only)
int launch_recover(const LaunchCtx &L, const KFields &F, const SharedTabs &S, int rec, bool heat,
                   hipStream_t s);
int launch_fix_recovery(const LaunchCtx &L, const KFields &F, double *const *flag, int *cluster,
                        hipStream_t s);
int launch_exchange(const LaunchCtx &L, const KFields &F, double *const *flag, const int *nbr,
                    hipStream_t s);
int launch_bc(const LaunchCtx &L, const KFields &F, double *const *flag, hipStream_t s);
// red: nblocks * interior planes * 3 doubles of scratch
int launch_dt(const LaunchCtx &L, double *const *alfven, double *red, double *state,
              double dt_factor, double *est_out, hipStream_t s);

}  // namespace pangu
