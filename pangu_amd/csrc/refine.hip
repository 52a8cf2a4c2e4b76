// Static mesh refinement on the GPU: ghost cells across refinement levels
// and flux correction (the pieces Parthenon adds to PANGU's task list on a
// multilevel mesh: AddBoundaryExchangeTasks and AddFluxCorrectionTasks,
// ideal_grmhd.cc:71-72,78).
//
// The host resolves the block tree into gather tables (refine_host.cpp); the
// kernels run one table entry per thread:
// - k_xfill: a ghost cell covered by a block of its own level copies that
//   block's interior cell, one covered by a finer block takes the mean of
//   the 2^dim fine cells (interior cells of the finer block);
// - k_xprol (after k_xfill): a ghost cell covered by a coarser block takes
//   the coarse value plus, per active dim, +-1/4 of the minmod-limited
//   difference to the coarse neighbours.  A coarse neighbour that is not a
//   coarse cell of a leaf is the mean of the fine block's own 2^dim cells
//   there (its interior or ghosts filled by k_xfill);
// - k_flux_correct: a coarse face flux on a face shared with finer blocks
//   becomes the mean of the 2^(dim-1) fine face fluxes.
// Sums run in a fixed tree order, so results match the oracle bit for bit.
// k_xfill runs one entry per thread over all components (its entries are
// cheap copies: re-reading the entry per component costs more than it
// hides); k_xprol and k_flux_correct run one entry per thread and component
// (blockIdx.y; the recovery flag is the last y index), so their seven and
// four gathers per value are independent across a wavefront.
#include <hip/hip_runtime.h>

#include "common.h"
#include "kernels.h"

namespace pangu {

__device__ __forceinline__ int block_dim(const DevBlock &B) {
  return B.n[2] > 1 ? 3 : (B.n[1] > 1 ? 2 : 1);
}

// Mean of the 2^dim cells from c (i fastest, then j, then k).
__device__ __forceinline__ double fine_mean(gptr<const double> a, int c, int dim, int sj, int sk) {
  if (dim == 1) return (a[c] + a[c + 1]) * 0.5;
  const double lo = (a[c] + a[c + 1]) + (a[c + sj] + a[c + sj + 1]);
  if (dim == 2) return lo * 0.25;
  const double hi = (a[c + sk] + a[c + sk + 1]) + (a[c + sk + sj] + a[c + sk + sj + 1]);
  return (lo + hi) * 0.125;
}

// Minimum of the same cells (the recovery flag: a coarse cell is good only
// if every fine cell is).
__device__ __forceinline__ double fine_min(gptr<const double> a, int c, int dim, int sj, int sk) {
  double m = fmin(a[c], a[c + 1]);
  if (dim == 1) return m;
  m = fmin(m, fmin(a[c + sj], a[c + sj + 1]));
  if (dim == 2) return m;
  return fmin(m, fmin(fmin(a[c + sk], a[c + sk + 1]), fmin(a[c + sk + sj], a[c + sk + sj + 1])));
}

__device__ __forceinline__ double minmod(double a, double b) {
  if (a > 0.0 && b > 0.0) return a < b ? a : b;
  if (a < 0.0 && b < 0.0) return a > b ? a : b;
  return 0.0;
}

__global__ __launch_bounds__(256) void k_xfill(KCtx C, KFields F, double *const *flag,
                                               const XFill *tab, int n) {
  const int e = (int)(blockIdx.x * 256 + threadIdx.x);
  if (e >= n) return;
  const XFill X = tab[e];
  const int np = C.np;
  if (!X.mean) {
    for (int q = 0; q < np; ++q) glb(F.prim[X.db * np + q])[X.dst] = glb(F.prim[X.sb * np + q])[X.src];
    if (flag) glb(flag[X.db])[X.dst] = glb(flag[X.sb])[X.src];
    return;
  }
  const DevBlock &S = C.blk[X.sb];
  const int dim = block_dim(S), sj = S.n[0], sk = S.n[0] * S.n[1];
  for (int q = 0; q < np; ++q)
    glb(F.prim[X.db * np + q])[X.dst] = fine_mean(glb(F.prim[X.sb * np + q]), X.src, dim, sj, sk);
  if (flag) glb(flag[X.db])[X.dst] = fine_min(glb(flag[X.sb]), X.src, dim, sj, sk);
}

__global__ __launch_bounds__(256) void k_xprol(KCtx C, KFields F, double *const *flag,
                                               const XProl *tab, int n) {
  const int e = (int)(blockIdx.x * 256 + threadIdx.x);
  if (e >= n) return;
  const XProl X = tab[e];
  const DevBlock &B = C.blk[X.b];
  const int dim = block_dim(B), sj = B.n[0], sk = B.n[0] * B.n[1];
  const int np = C.np, ns = 1 + 2 * dim, q = (int)blockIdx.y;
  if (q == np) {  // the flag row (only when flag != null): copy of the coarse value
    const double f = glb(flag[X.sb[0]])[X.sc[0]];
    const auto out = glb(flag[X.b]);
    for (int c = 0; c < (1 << dim); ++c)
      if ((X.mask >> c) & 1) out[X.base + (c & 1) + ((c >> 1) & 1) * sj + ((c >> 2) & 1) * sk] = f;
    return;
  }
  const auto own = glb(F.prim[X.b * np + q]);
  double v[7];
  for (int s = 0; s < ns; ++s)
    v[s] = X.sb[s] >= 0 ? glb(F.prim[X.sb[s] * np + q])[X.sc[s]] : fine_mean(own, X.sc[s], dim, sj, sk);
  double sl[3];
  for (int d = 0; d < dim; ++d) sl[d] = 0.25 * minmod(v[0] - v[1 + 2 * d], v[2 + 2 * d] - v[0]);
  for (int c = 0; c < (1 << dim); ++c) {
    if (!((X.mask >> c) & 1)) continue;
    double w = v[0];
    for (int d = 0; d < dim; ++d) w = ((c >> d) & 1) ? w + sl[d] : w - sl[d];
    own[X.base + (c & 1) + ((c >> 1) & 1) * sj + ((c >> 2) & 1) * sk] = w;
  }
}

__global__ __launch_bounds__(256) void k_flux_correct(KCtx C, KFields F, const XFlux *tab, int n) {
  const int e = (int)(blockIdx.x * 256 + threadIdx.x);
  if (e >= n) return;
  const XFlux X = tab[e];
  const DevBlock &B = C.blk[X.fb];
  const int dim = block_dim(B);
  const int stride[3] = {1, B.n[0], B.n[0] * B.n[1]};
  // tangential dims in increasing order
  const int a = stride[X.d == 0 ? 1 : 0], b = stride[X.d == 2 ? 1 : 2];
  const int nc = C.nc, c = X.fface, q = (int)blockIdx.y;
  const auto f = glb(F.flux[X.d][X.fb * nc + q]);
  double m;
  if (dim == 1)
    m = f[c];
  else if (dim == 2)
    m = (f[c] + f[c + a]) * 0.5;
  else
    m = ((f[c] + f[c + a]) + (f[c + b] + f[c + a + b])) * 0.25;
  glb(F.flux[X.d][X.cb * nc + q])[X.cface] = m;
}

int launch_exchange_ml(const LaunchCtx &L, const KFields &F, double *const *flag, hipStream_t s) {
  const MultiLevel &M = L.ml;
  if (M.nfill > 0) {
    hipLaunchKernelGGL(k_xfill, dim3((M.nfill + 255) / 256), dim3(256), 0, s, L.k, F, flag, M.fill,
                       M.nfill);
    if (L.clock) (void)L.clock->mark("exchange_fill", s);
  }
  if (M.nprol > 0) {
    hipLaunchKernelGGL(k_xprol, dim3((M.nprol + 255) / 256, L.k.np + (flag ? 1 : 0)), dim3(256), 0, s, L.k,
                       F, flag, M.prol, M.nprol);
    if (L.clock) (void)L.clock->mark("exchange_prolong", s);
  }
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_flux_correction(const LaunchCtx &L, const KFields &F, hipStream_t s) {
  const MultiLevel &M = L.ml;
  if (M.nflux > 0) {
    hipLaunchKernelGGL(k_flux_correct, dim3((M.nflux + 255) / 256, L.k.nc), dim3(256), 0, s, L.k, F, M.flux,
                       M.nflux);
    if (L.clock) (void)L.clock->mark("flux_correction", s);
  }
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // namespace pangu
