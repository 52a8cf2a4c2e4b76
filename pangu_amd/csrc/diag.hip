// Run diagnostics on the GPU: shell integrals of the mass, energy, angular
// momentum and magnetic fluxes through a surface x1 = const (the quantities
// of doc/grmhd_intro.md:556-574 in code coordinates), volume integrals of
// the conserved mass and energy, and the CT invariant (corner div B).  The
// reference computes none of these itself.
//
// Reductions run in a fixed order (a tree per 256-thread tile, then one
// workgroup per mesh over the tiles), so results repeat bit for bit from
// run to run.  On a Cartesian Kerr-Schild mesh, cells inside the excision
// radius (those Recovery skips, invertor.cc:83-91) hold no evolved state:
// their U can be non-finite after an update.  They are left out of the
// volume integrals and |U_B|, and corners that touch one are left out of div B.
#include <hip/hip_runtime.h>

#include <mutex>

#include "common.h"
#include "kernels.h"
#include "physics_dev.h"
#include "tiles_dev.h"

namespace pangu {

// Workspace of pangu_diagnostics_noalloc: the per-tile partial sums, then
// the per-mesh tile ranges and shell indices as ints.  One copy per device,
// held until the process ends; every call on a device shares it.  Host code
// never passes `scratch` itself to a kernel (on the host the name is only a
// stand-in for the device symbol): it uses the address that
// diag_scratch_address() looks up once per device.
static __device__ __align__(256) double scratch[kDiagScratchDoubles];

namespace {
std::mutex g_scratch_mu;
double *g_scratch_ptr[kDiagMaxDevices] = {};
}  // namespace

hipError_t diag_scratch_address(int device, double **out) {
  *out = nullptr;
  if (device < 0 || device >= kDiagMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return e;
    if (cur != device) return hipErrorInvalidDevice;  // the caller makes `device` current
    double *ptr = nullptr;
    e = hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[device] = ptr;
  }
  *out = g_scratch_ptr[device];
  return hipSuccess;
}

template <bool HB>
__global__ __launch_bounds__(256) void k_diag(KCtx C, KFields F, const int *shell, double *part,
                                              Range3 R) {
  __shared__ double red[PANGU_NDIAG][TILE];
  const int N = R.ntiles;
  const int n8 = (N + 7) >> 3;
  const int L = (int)blockIdx.x;
  const int T = (L & 7) * n8 + (L >> 3);
  if (T >= N) return;  // uniform over the workgroup
  double v[PANGU_NDIAG];
#pragma unroll
  for (int q = 0; q < PANGU_NDIAG; ++q) v[q] = 0.0;
  int b, k, j, i;
  if (decode_tile(C, R, b, k, j, i)) {
    const DevBlock &B = C.blk[b];
    const DevGeo &G = C.geo[B.geo];
    const int nc = C.nc;
    const int64_t sj = B.n[0], sk = (int64_t)B.n[0] * B.n[1];
    const int64_t id = k * sk + j * sj + i;
    const double dA = B.dx[1] * B.dx[2];
    const double dV = B.dx[0] * B.dx[1] * B.dx[2];
    const bool cks = C.ph.metric == PANGU_METRIC_CKS;
    auto excised = [&](int kk, int jj, int ii) {
      const int64_t t = ((int64_t)(G.has3d ? kk : 0) * G.nj + jj) * G.ni + ii;
      return cks && glb(G.cen)[CG_RAD * G.ncen + t] < C.ph.r_excise;
    };
    const bool exc = excised(k, j, i);
    if (!exc) {
      v[PANGU_DIAG_MASS] = glb(F.cons[b * nc])[id] * dV;
      v[PANGU_DIAG_ENERGY] = glb(F.cons[b * nc + 1])[id] * dV;
    }
    const int nx0 = B.e[0] - B.s[0] + 1;
    if (B.lx[0] * nx0 + (i - B.s[0]) == shell[B.mesh]) {
      const int64_t t = ((int64_t)(G.has3d ? k : 0) * G.nj + j) * G.ni + i;
      const auto c = glb(G.cen);
      const int64_t s = G.ncen;
      double g[16], gc0[4];
#pragma unroll
      for (int n = 0; n < 16; ++n) g[n] = c[(CG_GCOV + n) * s + t];
#pragma unroll
      for (int n = 0; n < 4; ++n) gc0[n] = c[(CG_GCON + n) * s + t];
      const double alpha = c[CG_ALPHA * s + t], sg = c[CG_SG * s + t];
      double p[NPRIM];
#pragma unroll
      for (int n = 0; n < NPRIM; ++n) p[n] = 0.0;
      for (int q = 0; q < 5; ++q) p[q] = glb(F.prim[b * C.np + q])[id];
      if (C.iBX >= 0)
        for (int q = 0; q < 3; ++q) p[BX1 + q] = glb(F.prim[b * C.np + C.iBX + q])[id];
      St st;
      calc_state(p, g, gc0, alpha, st);
      const double gam = C.ph.gamma;
      const double pg = (gam - 1.) * p[ENY];
      const double h = p[RHO] + gam * p[ENY];
      const double etot = h + st.bsq, ptot = pg + 0.5 * st.bsq;
      const double t0 = tmunu<1>(st, etot, ptot, 0), t3 = tmunu<1>(st, etot, ptot, 3);
      v[PANGU_DIAG_MDOT] = -(p[RHO] * st.ucon[1] * sg) * dA;
      v[PANGU_DIAG_EDOT] = t0 * sg * dA;
      v[PANGU_DIAG_JDOT] = -(t3 * sg) * dA;
      v[PANGU_DIAG_PHI] = 0.5 * fabs(p[BX1]) * sg * dA;
    }
    if (HB) {
      const int ndim = (B.n[0] > 1) + (B.n[1] > 1) + (B.n[2] > 1);
      gptr<double> ub[3];
      for (int d = 0; d < 3; ++d) ub[d] = glb(F.cons[b * nc + C.iBX + d]);
      double m = 0.0;
      if (!exc)
        for (int d = 0; d < ndim; ++d) m = fmax(m, fabs(ub[d][id]));
      v[PANGU_DIAG_UBMAX] = m;
      // whether the cells around corner (k, j, i) - 1/2 include an excised
      // one (called only where those cells exist: i > s0, j > s1, and k > s2 in 3-D)
      auto touch = [&](int nk) {
        bool t = false;
        if (cks)
          for (int dk = 0; dk < nk; ++dk)
            for (int dj = 0; dj < 2; ++dj)
              for (int di = 0; di < 2; ++di) t = t || excised(k - dk, j - dj, i - di);
        return t;
      };
      if (ndim == 2 && i > B.s[0] && j > B.s[1] && !touch(1)) {
        const auto b1 = ub[0], b2 = ub[1];
        const double d1 = (b1[id] + b1[id - sj] - b1[id - 1] - b1[id - sj - 1]) / (2 * B.dx[0]);
        const double d2 = (b2[id] + b2[id - 1] - b2[id - sj] - b2[id - sj - 1]) / (2 * B.dx[1]);
        v[PANGU_DIAG_DIVB] = fabs(d1 + d2);
      } else if (ndim == 3 && i > B.s[0] && j > B.s[1] && k > B.s[2] && !touch(2)) {
        double tot = 0.0;
        for (int d = 0; d < 3; ++d) {
          double acc = 0.0;
          for (int dk = 0; dk < 2; ++dk)
            for (int dj = 0; dj < 2; ++dj)
              for (int di = 0; di < 2; ++di) {
                const int off = d == 0 ? di : (d == 1 ? dj : dk);
                const double sign = off == 0 ? 1.0 : -1.0;
                acc = acc + sign * ub[d][id - dk * sk - dj * sj - di];
              }
          tot = tot + acc / (4 * B.dx[d]);
        }
        v[PANGU_DIAG_DIVB] = fabs(tot);
      }
    }
  }
  const int tid = threadIdx.x;
#pragma unroll
  for (int q = 0; q < PANGU_NDIAG; ++q) red[q][tid] = v[q];
  __syncthreads();
  for (int w = TILE / 2; w > 0; w >>= 1) {
    if (tid < w) {
#pragma unroll
      for (int q = 0; q < PANGU_NDIAG; ++q)
        red[q][tid] = (q < PANGU_DIAG_DIVB) ? red[q][tid] + red[q][tid + w]
                                            : fmax(red[q][tid], red[q][tid + w]);
    }
    __syncthreads();
  }
  if (tid < PANGU_NDIAG) glb(part)[(int64_t)T * PANGU_NDIAG + tid] = red[tid][0];
}

// One workgroup per mesh: the tiles of mesh m are [mtile[m], mtile[m + 1]).
__global__ __launch_bounds__(256) void k_diag_final(const double *part, const int *mtile, double *out) {
  __shared__ double red[PANGU_NDIAG][TILE];
  const int m = blockIdx.x, tid = threadIdx.x;
  const int t0 = mtile[m], t1 = mtile[m + 1];
  double v[PANGU_NDIAG];
#pragma unroll
  for (int q = 0; q < PANGU_NDIAG; ++q) v[q] = 0.0;
  for (int T = t0 + tid; T < t1; T += TILE)
#pragma unroll
    for (int q = 0; q < PANGU_NDIAG; ++q) {
      const double x = glb(part)[(int64_t)T * PANGU_NDIAG + q];
      v[q] = (q < PANGU_DIAG_DIVB) ? v[q] + x : fmax(v[q], x);
    }
#pragma unroll
  for (int q = 0; q < PANGU_NDIAG; ++q) red[q][tid] = v[q];
  __syncthreads();
  for (int w = TILE / 2; w > 0; w >>= 1) {
    if (tid < w) {
#pragma unroll
      for (int q = 0; q < PANGU_NDIAG; ++q)
        red[q][tid] = (q < PANGU_DIAG_DIVB) ? red[q][tid] + red[q][tid + w]
                                            : fmax(red[q][tid], red[q][tid + w]);
    }
    __syncthreads();
  }
  if (tid < PANGU_NDIAG) glb(out)[(int64_t)m * PANGU_NDIAG + tid] = red[tid][0];
}

int launch_diagnostics(const LaunchCtx &L, const KFields &F, const int *shell, double *part,
                       const int *mtile, int nmesh, double *out, hipStream_t s) {
  const Range3 &R = L.interior;
  if (R.ntiles > 0) {
    const dim3 g((unsigned)(((R.ntiles + 7) / 8) * 8));
    if (L.k.iBX >= 0)
      hipLaunchKernelGGL((k_diag<true>), g, dim3(TILE), 0, s, L.k, F, shell, part, R);
    else
      hipLaunchKernelGGL((k_diag<false>), g, dim3(TILE), 0, s, L.k, F, shell, part, R);
  }
  hipLaunchKernelGGL(k_diag_final, dim3((unsigned)nmesh), dim3(TILE), 0, s, part, mtile, out);
  return (int)hipGetLastError();
}

}  // namespace pangu
