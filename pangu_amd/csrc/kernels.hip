// gfx950 kernels of the GRMHD per-RK-stage update path (FP64 vector ALU).
//
// Thread mapping: a launch is a 1D grid of 256-thread tiles.  The (i, j)
// range of each (block, k) plane is flattened row-major, so consecutive lanes
// take consecutive i (coalesced loads) and rows of any width (65, 66, 72 ...)
// fill whole wavefronts.  All three flux sweeps therefore load coalesced
// rows; the reference's j- and k-innermost sweeps (laxf.cc:189,295) would
// stride through memory on a GPU.  Tiles are remapped so that consecutive
// logical tiles run on the same XCD: the hardware deals workgroups to the 8
// XCDs round-robin, and neighbouring tiles share stencil rows and planes
// through that XCD's L2.
#include <hip/hip_runtime.h>
#include <float.h>

#include "common.h"
#include "kernels.h"
#include "physics_dev.h"
#include "recovery_dev.h"
#include "tiles_dev.h"

namespace pangu {

// Packed component q -> slot, for the enabled field set
// (package_registration.cc:84-96).
template <bool HB, bool HK>
__device__ __forceinline__ constexpr int slot_of(int q) {
  return q < 5 ? q : (HB ? (q < 8 ? q : (q == 8 ? ENT : KEL)) : (q == 5 ? ENT : KEL));
}
template <bool HB, bool HK>
__device__ __forceinline__ constexpr int ncomp() {
  return 6 + (HB ? 3 : 0) + (HK ? 1 : 0);
}
template <bool HB>
__device__ __forceinline__ constexpr int ient() {
  return HB ? 8 : 5;
}
template <bool HB>
__device__ __forceinline__ constexpr int ikel() {
  return HB ? 9 : 6;
}

__device__ __forceinline__ int64_t cen_index(const DevGeo &G, int k, int j, int i) {
  return ((int64_t)(G.has3d ? k : 0) * G.nj + j) * G.ni + i;
}

// Loads of the centre table.
struct CenLoad {
  double g[16], gc0[4], alpha, sg;
};
__device__ __forceinline__ void load_cen_state(const DevGeo &G, int64_t t, CenLoad &m) {
  const double *c = G.cen;
  const int64_t s = G.ncen;
  const uint32_t t32 = (uint32_t)t;
#pragma unroll
  for (int n = 0; n < 16; ++n) m.g[n] = ldc32(c, (CG_GCOV + n) * s, t32);
#pragma unroll
  for (int n = 0; n < 4; ++n) m.gc0[n] = ldc32(c, (CG_GCON + n) * s, t32);
  m.alpha = ldc32(c, CG_ALPHA * s, t32);
  m.sg = ldc32(c, CG_SG * s, t32);
}

// ------------------------------------------------ fix_primitive + conservative
// FixPrimitive over the entire block (primitive_fixer.cc:64-166) and, for
// interior cells, CalculateConservative (conservative.cc:57-107).  The
// primitives and conservatives are written with non-temporal stores: on
// MI355X that leaves L2 to the stencil reads of the flux sweep that follows
// (measured 1 % faster per cycle on the 3-D torus than ordinary stores).
// One cell of k_fix_cons.  write = false computes without storing (a halo
// cell of k_fix_flux_x1, which another workgroup owns); pout (optional)
// receives the fixed primitives, every slot.
template <bool FIX, bool CONS, bool HB, bool HK>
__device__ __forceinline__ void fix_cons_cell(const KCtx &C, const KFields &F, uint8_t *const *mask, int b,
                                              int k, int j, int i, bool write, double *pout) {
  const DevBlock &B = C.blk[b];
  const bool interior = i >= B.s[0] && i <= B.e[0] && j >= B.s[1] && j <= B.e[1] &&
                        k >= B.s[2] && k <= B.e[2];
  if (!FIX && !interior) return;
  const DevGeo &G = C.geo[B.geo];
  const int64_t t = cen_index(G, k, j, i);
  const int64_t id = ((int64_t)k * B.n[1] + j) * B.n[0] + i;
  constexpr int NC = ncomp<HB, HK>();
  gptr<double> pp[NC];
#pragma unroll
  for (int q = 0; q < NC; ++q) pp[q] = glb(F.prim[b * NC + q]);
  double p[NPRIM];
#pragma unroll
  for (int s = 0; s < NPRIM; ++s) p[s] = 0.0;
#pragma unroll
  for (int q = 0; q < NC; ++q) p[slot_of<HB, HK>(q)] = pp[q][id];
  CenLoad m;
  load_cen_state(G, t, m);
  const Phys &P = C.ph;
  if (FIX) {
    const double small = 1.0e-20;
    const double sms = kmax(P.sigma_max, small);
    const double rho_before = kmax(p[RHO], small);
    double rho_floor = ldc32(G.cen, (int64_t)CG_RHOFL * G.ncen, (uint32_t)t);
    const double eng_floor = ldc32(G.cen, (int64_t)CG_ENGFL * G.ncen, (uint32_t)t);
    St st;
    calc_state(p, m.g, m.gc0, m.alpha, st);
    const double rho = p[RHO], ug = p[ENY];
    uint8_t mk = 0;
    if (rho_floor < ug / sms) rho_floor = ug / sms;
    const double sigma = st.bsq / rho;
    if (sigma > P.sigma_max) {
      rho_floor = st.bsq / sms;
      mk |= 8;
    }
    if (p[RHO] < rho_floor) {
      p[RHO] = rho_floor;
      mk |= 1;
    }
    if (p[ENY] < eng_floor) {
      p[ENY] = eng_floor;
      mk |= 2;
    }
    const double u1 = p[UX1], u2 = p[UX2], u3 = p[UX3];
    const double *g = m.g;
    const double uvsq = g[5] * u1 * u1 + g[10] * u2 * u2 + g[15] * u3 * u3 +
                        2.0 * (g[6] * u1 * u2 + g[7] * u1 * u3 + g[11] * u2 * u3);
    const double umax2 = P.lorentz_max * P.lorentz_max - 1.0;
    if (umax2 > 0.0 && uvsq > umax2) {
      const double f = sqrt(umax2 / kmax(uvsq, small));
      p[UX1] *= f;
      p[UX2] *= f;
      p[UX3] *= f;
      mk |= 4;
    }
    const double rho_after = kmax(p[RHO], small);
    if (HK && rho_after != rho_before) p[KEL] *= pow(rho_after / rho_before, -P.gamma);
    p[ENT] = (P.gamma - 1.0) * p[ENY] * pow(p[RHO], -P.gamma);
    if (HK) p[KEL] = clamp_by_ratio(P.ratio_min, P.ratio_max, p[ENT], p[KEL]);
    if (P.metric == PANGU_METRIC_CKS && ldc32(G.cen, (int64_t)CG_RAD * G.ncen, (uint32_t)t) < P.r_excise) {
      const double e_exc = P.pexcise / (P.gamma - 1.0);
      p[RHO] = P.dexcise;
      p[ENY] = e_exc;
      p[UX1] = 0.0;
      p[UX2] = 0.0;
      p[UX3] = 0.0;
      p[ENT] = (P.gamma - 1.0) * e_exc * pow(P.dexcise, -P.gamma);
      if (HK) p[KEL] = e_exc;
      mk |= 16;
    }
    // rho, u and u~ change only where a floor, the Lorentz cap or excision
    // set a mask bit; elsewhere memory already holds them (the entropy is
    // recomputed in every cell)
    if (write && mk) {
      __builtin_nontemporal_store(p[RHO], &pp[0][id]);
      __builtin_nontemporal_store(p[ENY], &pp[1][id]);
      __builtin_nontemporal_store(p[UX1], &pp[2][id]);
      __builtin_nontemporal_store(p[UX2], &pp[3][id]);
      __builtin_nontemporal_store(p[UX3], &pp[4][id]);
    }
    if (write) {
      __builtin_nontemporal_store(p[ENT], &pp[ient<HB>()][id]);
      if (HK) pp[ikel<HB>()][id] = p[KEL];
      if (mask) glb(mask[b])[id] = mk;
    }
  }
  if (pout) {
#pragma unroll
    for (int q = 0; q < NPRIM; ++q) pout[q] = p[q];
  }
  if (CONS && interior && write) {
    St st;
    calc_state(p, m.g, m.gc0, m.alpha, st);
    double f[NPRIM];
    contravariant_flux<0>(P.gamma, p, st, m.sg, f);
#pragma unroll
    for (int q = 0; q < NC; ++q) __builtin_nontemporal_store(f[slot_of<HB, HK>(q)], &glb(F.cons[b * NC + q])[id]);
  }
}

template <bool FIX, bool CONS, bool HB, bool HK>
__global__ __launch_bounds__(256) void k_fix_cons(KCtx C, KFields F, uint8_t *const *mask, Range3 R) {
  int b, k, j, i;
  if (!decode_tile(C, R, b, k, j, i)) return;
  fix_cons_cell<FIX, CONS, HB, HK>(C, F, mask, b, k, j, i, true, nullptr);
}

// ------------------------------------------------------------- flux sweep
// One direction of CalculateLAXF / CalculateHLL (laxf.cc:78-181 and the
// X2/X3 twins): reconstruction of every primitive, face metric, fast speeds,
// four contravariant fluxes and the Riemann combine.  Face (k,j,i) is the
// lower face of cell (k,j,i) along DIR.
// Face physics of one sweep direction for face (k, j, i) of block b, the
// lower face of cell (k, j, i) along DIR: face metric, fluid states, fast
// speeds, the four contravariant fluxes and the Riemann combine, from the
// reconstructed left (pl) and right (pr) primitives.
template <int DIR, int RS, bool HB, bool HK>
__device__ __forceinline__ void face_flux(const KCtx &C, const KFields &F, double *const *alfven, int b,
                                          int k, int j, int i, const double pl[NPRIM],
                                          const double pr[NPRIM]) {
  const DevBlock &B = C.blk[b];
  const int64_t id = ((int64_t)k * B.n[1] + j) * B.n[0] + i;
  constexpr int NC = ncomp<HB, HK>();
  const DevGeo &G = C.geo[B.geo];
  double g[16], gc0[4], gdd, sg, alpha, valid;
  if (DIR == 3 && !G.face[2]) {
    const uint32_t t = (uint32_t)cen_index(G, k, j, i);
    const double *c = G.cen;
    const uint32_t s = (uint32_t)G.ncen;
#pragma unroll
    for (int n = 0; n < 16; ++n) g[n] = ldc32(c, (int64_t)(CG_GCOV + n) * s, t);
#pragma unroll
    for (int n = 0; n < 4; ++n) gc0[n] = ldc32(c, (int64_t)(CG_GCON + n) * s, t);
    gdd = ldc32(c, (int64_t)(CG_GCON + 15) * s, t);
    sg = ldc32(c, (int64_t)CG_SG * s, t);
    alpha = ldc32(c, (int64_t)CG_ALPHA * s, t);
    valid = 1.0;
  } else {
    const int d = DIR - 1;
    const uint32_t t = (uint32_t)(((G.has3d ? k : 0) * G.fj[d] + j) * G.fi[d] + i);
    const double *c = G.face[d];
    const uint32_t s = (uint32_t)G.nface[d];
#pragma unroll
    for (int n = 0; n < 16; ++n) g[n] = ldc32(c, (int64_t)(FG_GCOV + n) * s, t);
#pragma unroll
    for (int n = 0; n < 4; ++n) gc0[n] = ldc32(c, (int64_t)(FG_GCON0 + n) * s, t);
    gdd = ldc32(c, (int64_t)FG_GCONDD * s, t);
    sg = ldc32(c, (int64_t)FG_SG * s, t);
    alpha = ldc32(c, (int64_t)FG_ALPHA * s, t);
    valid = ldc32(c, (int64_t)FG_VALID * s, t);
  }
  double *const al = alfven[b * 3 + (DIR - 1)];
  const uint32_t id32 = (uint32_t)id;
  if (valid == 0.0) {
    // singular face metric (polar axis): zero flux and wave speed
    st32(al, id32, 0.0);
#pragma unroll
    for (int q = 0; q < NC; ++q) st32(F.flux[DIR - 1][b * NC + q], id32, 0.0);
    return;
  }
  const double gam = C.ph.gamma;
  St sl, sr;
  calc_state(pl, g, gc0, alpha, sl);
  calc_state(pr, g, gc0, alpha, sr);
  double vmaxL, vminL, vmaxR, vminR;
  fast_ms<DIR>(gam, pl, sl, gc0[0], gc0[DIR], gdd, vmaxL, vminL);
  fast_ms<DIR>(gam, pr, sr, gc0[0], gc0[DIR], gdd, vmaxR, vminR);
  const double cmax = fabs(kmax(kmax(0., vmaxL), vmaxR));
  const double cmin = fabs(kmax(kmax(0., -vminL), -vminR));
  const double cc = kmax(cmax, cmin);
  st32(al, id32, cc);
  double UL[NPRIM], UR[NPRIM], FL[NPRIM], FR[NPRIM];
  contravariant_flux<0>(gam, pl, sl, sg, UL);
  contravariant_flux<0>(gam, pr, sr, sg, UR);
  contravariant_flux<DIR>(gam, pl, sl, sg, FL);
  contravariant_flux<DIR>(gam, pr, sr, sg, FR);
  const double den = cmax + cmin + 1e-20, rden = 1.0 / den;
#pragma unroll
  for (int q = 0; q < NC; ++q) {
    const int s = slot_of<HB, HK>(q);
    double f;
    if (RS == PANGU_RIEMANN_LAXF)
      f = 0.5 * (FL[s] + FR[s] - cc * (UR[s] - UL[s]));
    else
      f = div_by(cmax * FL[s] + cmin * FR[s] - cmax * cmin * (UR[s] - UL[s]), den, rden);
    st32(F.flux[DIR - 1][b * NC + q], id32, f);
  }
}


template <int DIR, int LIM, int RS, bool HB, bool HK>
__global__ __launch_bounds__(256) void k_flux(KCtx C, KFields F, double *const *alfven, Range3 R) {
  int b, k, j, i;
  if (!decode_tile(C, R, b, k, j, i)) return;
  const DevBlock &B = C.blk[b];
  const int64_t sj = B.n[0], sk = (int64_t)B.n[0] * B.n[1];
  const int64_t st = (DIR == 1) ? 1 : (DIR == 2 ? sj : sk);
  const int64_t id = k * sk + j * sj + i;
  constexpr int NC = ncomp<HB, HK>();
  double pl[NPRIM], pr[NPRIM];
#pragma unroll
  for (int s = 0; s < NPRIM; ++s) {
    pl[s] = 0.0;
    pr[s] = 0.0;
  }
  // Every stencil value of every variable is loaded before the first
  // reconstruction, so a wavefront waits for memory once rather than once
  // per variable (the compiler keeps the loads of an unrolled per-variable
  // loop behind each variable's arithmetic).  The peak register use stays in
  // the face physics.
  constexpr int NS = (LIM == PANGU_LIMITER_PPM4) ? 6 : 4;  // offsets -NS/2 .. NS/2 - 1
  double v[NC][NS];
#pragma unroll
  for (int q = 0; q < NC; ++q) {
    const double *a = F.prim[b * NC + q];
#pragma unroll
    for (int n = 0; n < NS; ++n)
      // x1 neighbours fold into the load's immediate offset; x2/x3 offsets
      // (a row, a plane) into the scalar base
      v[q][n] = DIR == 1 ? glb(a)[id + (n - NS / 2)] : ld32(a, (n - NS / 2) * st, (uint32_t)id);
  }
#pragma unroll
  for (int q = 0; q < NC; ++q) {
    const int s = slot_of<HB, HK>(q);
    double t0, t1;
    if (LIM == PANGU_LIMITER_PPM4) {
      recon_ppm4(v[q][0], v[q][1], v[q][2], v[q][3], v[q][4], pl[s], t0);
      recon_ppm4(v[q][1], v[q][2], v[q][3], v[q][4], v[q][5], t1, pr[s]);
    } else {
      recon_mc(v[q][0], v[q][1], v[q][2], pl[s], t0);
      recon_mc(v[q][1], v[q][2], v[q][3], t1, pr[s]);
    }
  }
  face_flux<DIR, RS, HB, HK>(C, F, alfven, b, k, j, i, pl, pr);
}

// FixPrimitive + CalculateConservative (k_fix_cons) and the X1 sweep
// (k_flux<1>) in one pass.  A workgroup of 256 threads covers cells
// t0 - 4 .. t0 + R.tile + 3 of one ghosted plane (R: the entire range in runs
// of R.tile = 248 cells, decoded as decode_tile_at does), fixes every one of
// them and stores the outputs of the cells it owns (t0 .. t0 + 247); the
// four halo cells on each side are fixed again by the neighbouring
// workgroup (FixPrimitive is pointwise, so both compute the same bits).  The
// fixed primitives wait in LDS, and each owned cell that is an X1 face
// (s..e+1, one transverse ghost layer) reconstructs its two states from
// them: the stencil i-3..i+2 lies in the cell's row and in the window.  Same
// operations and operands as the two kernels: identical results.
template <int LIM, int RS, bool HB, bool HK>
__global__ __launch_bounds__(256) void k_fix_flux_x1(KCtx C, KFields F, uint8_t *const *mask,
                                                     double *const *alfven, Range3 R) {
  __shared__ double sp[NPRIM][256];
  const int N = R.ntiles;
  const int n8 = (N + 7) >> 3;
  const int L = (int)blockIdx.x;
  const int T = (L & 7) * n8 + (L >> 3);
  if (T >= N) return;  // uniform over the workgroup
  const int TL = R.tile;
  int lo[3], cnt[3], local, tpp, b;
  if (R.uni) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      lo[d] = R.ulo[d];
      cnt[d] = R.ucnt[d];
    }
    tpp = R.utpp;
    const int per_block = cnt[2] * tpp;
    b = T / per_block;
    local = T - b * per_block;
  } else {
    b = R.tblock[T];
    local = T - R.tfirst[b];
    block_range(C.blk[b], R, lo, cnt);
    tpp = (cnt[0] * cnt[1] + TL - 1) / TL;
  }
  const int plane = cnt[0] * cnt[1];
  int tile, kk;
  if (R.order) {
    tile = local / cnt[2];
    kk = local - tile * cnt[2];
  } else {
    kk = local / tpp;
    tile = local - kk * tpp;
  }
  const int x = (int)threadIdx.x;
  const int t = tile * TL + x - 4;
  const bool valid = t >= 0 && t < plane && x < TL + 8;
  const bool own = valid && x >= 4 && x < TL + 4;
  const int k = lo[2] + kk;
  int j = 0, i = 0;
  if (valid) {
    const int jj = t / cnt[0];
    j = lo[1] + jj;
    i = lo[0] + (t - jj * cnt[0]);
  }
  double p[NPRIM];
#pragma unroll
  for (int q = 0; q < NPRIM; ++q) p[q] = 0.0;
  if (valid) fix_cons_cell<true, true, HB, HK>(C, F, mask, b, k, j, i, own, p);
#pragma unroll
  for (int q = 0; q < NPRIM; ++q) sp[q][x] = p[q];
  __syncthreads();
  if (!own) return;
  const DevBlock &B = C.blk[b];
  const bool face = i >= B.s[0] && i <= B.e[0] + 1 &&
                    (B.n[1] > 1 ? (j >= B.s[1] - 1 && j <= B.e[1] + 1) : j == B.s[1]) &&
                    (B.n[2] > 1 ? (k >= B.s[2] - 1 && k <= B.e[2] + 1) : k == B.s[2]);
  if (!face) return;
  constexpr int NC = ncomp<HB, HK>();
  constexpr int NS = (LIM == PANGU_LIMITER_PPM4) ? 6 : 4;  // offsets -NS/2 .. NS/2 - 1
  double pl[NPRIM], pr[NPRIM];
#pragma unroll
  for (int q = 0; q < NPRIM; ++q) {
    pl[q] = 0.0;
    pr[q] = 0.0;
  }
#pragma unroll
  for (int q = 0; q < NC; ++q) {
    const int sl = slot_of<HB, HK>(q);
    double v[NS];
#pragma unroll
    for (int n = 0; n < NS; ++n) v[n] = sp[sl][x + n - NS / 2];
    double t0, t1;
    if (LIM == PANGU_LIMITER_PPM4) {
      recon_ppm4(v[0], v[1], v[2], v[3], v[4], pl[sl], t0);
      recon_ppm4(v[1], v[2], v[3], v[4], v[5], t1, pr[sl]);
    } else {
      recon_mc(v[0], v[1], v[2], pl[sl], t0);
      recon_mc(v[1], v[2], v[3], t1, pr[sl]);
    }
  }
  face_flux<1, RS, HB, HK>(C, F, alfven, b, k, j, i, pl, pr);
}

// The value of lane l - 1 (lane 0 gets 0) through the DPP wavefront shift
// (wave_shr:1), 32 bits at a time.
__device__ __forceinline__ double wave_shr1(double x) {
  const uint64_t u = (uint64_t)__double_as_longlong(x);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, 0x138, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32), 0x138, 0xf, 0xf, false);
  return __longlong_as_double((long long)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo));
}

// Exchange with the other lane of an aligned lane pair (l ^ 1) through the
// data-parallel-primitive path (quad_perm [1, 0, 3, 2]), 32 bits at a time.
__device__ __forceinline__ double pair_swap(double x) {
  const uint64_t u = __builtin_bit_cast(uint64_t, x);
  const int lo = __builtin_amdgcn_mov_dpp((int)(uint32_t)u, 0xB1, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp((int)(uint32_t)(u >> 32), 0xB1, 0xF, 0xF, false);
  return __builtin_bit_cast(double, ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

// Flux sweep with two lanes per face.  The even lane of a pair builds the
// left state (cell i - 1 along DIR), the odd lane the right state (cell i):
// each reconstructs one side (the stencil shared through the pair: each lane
// loads every other stencil cell and takes the rest from its partner), forms
// its fluid state, wave speeds and flux vectors, then the pair exchanges the
// speeds and the flux components the other lane combines; the even lane
// writes the first half of the components and the wave speed, the odd lane
// the rest.  Every value is computed by the same expressions in the same
// order as k_flux, so the results are identical; what changes is that one
// lane holds one fluid state instead of two (fewer registers, more
// wavefronts in flight).
template <int DIR, int LIM, int RS, bool HB, bool HK>
__global__ __launch_bounds__(256) void k_flux_pair(KCtx C, KFields F, double *const *alfven, Range3 R) {
  int b = 0, k = 0, j = 0, i = 0;
  if (!decode_tile_pair(C, R, b, k, j, i)) return;  // both lanes of a pair share the outcome
  const int side = (int)threadIdx.x & 1;            // 0: left state, 1: right state
  const DevBlock &B = C.blk[b];
  const int64_t sj = B.n[0], sk = (int64_t)B.n[0] * B.n[1];
  const int64_t st = (DIR == 1) ? 1 : (DIR == 2 ? sj : sk);
  const int64_t id = k * sk + j * sj + i;
  constexpr int NC = ncomp<HB, HK>();
  constexpr int NS = (LIM == PANGU_LIMITER_PPM4) ? 6 : 4;  // stencil offsets -NS/2 .. NS/2 - 1
  constexpr int NL = NS / 2;                               // cells each lane loads
  double p[NPRIM];
#pragma unroll
  for (int q = 0; q < NPRIM; ++q) p[q] = 0.0;
  {
    // lane `side` loads offsets -NS/2 + side + 2m; its stencil for one side
    // is offsets side - NS/2 + n, n = 0 .. NS - 2
    double a[NC][NL];
#pragma unroll
    for (int q = 0; q < NC; ++q) {
      const auto arr = glb(F.prim[b * NC + q]);
#pragma unroll
      for (int m = 0; m < NL; ++m) a[q][m] = arr[id + (2 * m + side - NS / 2) * st];
    }
#pragma unroll
    for (int q = 0; q < NC; ++q) {
      double x[NL], w[NS - 1];
#pragma unroll
      for (int m = 0; m < NL; ++m) x[m] = pair_swap(a[q][m]);
#pragma unroll
      for (int n = 0; n < NS - 1; ++n) {
        if ((n & 1) == 0)
          w[n] = a[q][n >> 1];
        else
          w[n] = side ? x[((n - 1) >> 1) + 1] : x[(n - 1) >> 1];
      }
      const int s = slot_of<HB, HK>(q);
      double t;
      if (LIM == PANGU_LIMITER_PPM4) {
        if (side == 0)
          recon_ppm4(w[0], w[1], w[2], w[3], w[4], p[s], t);
        else
          recon_ppm4(w[0], w[1], w[2], w[3], w[4], t, p[s]);
      } else {
        if (side == 0)
          recon_mc(w[0], w[1], w[2], p[s], t);
        else
          recon_mc(w[0], w[1], w[2], t, p[s]);
      }
    }
  }
  const DevGeo &G = C.geo[B.geo];
  double g[16], gc0[4], gdd, sg, alpha, valid;
  if (DIR == 3 && !G.face[2]) {
    const int64_t t = cen_index(G, k, j, i);
    const auto c = glb(G.cen);
    const int64_t ns = G.ncen;
#pragma unroll
    for (int n = 0; n < 16; ++n) g[n] = c[(CG_GCOV + n) * ns + t];
#pragma unroll
    for (int n = 0; n < 4; ++n) gc0[n] = c[(CG_GCON + n) * ns + t];
    gdd = c[(CG_GCON + 15) * ns + t];
    sg = c[CG_SG * ns + t];
    alpha = c[CG_ALPHA * ns + t];
    valid = 1.0;
  } else {
    const int d = DIR - 1;
    const int64_t t = ((int64_t)(G.has3d ? k : 0) * G.fj[d] + j) * G.fi[d] + i;
    const auto c = glb(G.face[d]);
    const int64_t ns = G.nface[d];
#pragma unroll
    for (int n = 0; n < 16; ++n) g[n] = c[(FG_GCOV + n) * ns + t];
#pragma unroll
    for (int n = 0; n < 4; ++n) gc0[n] = c[(FG_GCON0 + n) * ns + t];
    gdd = c[FG_GCONDD * ns + t];
    sg = c[FG_SG * ns + t];
    alpha = c[FG_ALPHA * ns + t];
    valid = c[FG_VALID * ns + t];
  }
  constexpr int NH = (NC + 1) / 2;  // components written by the even lane
  const auto al = glb(alfven[b * 3 + (DIR - 1)]);
  if (valid == 0.0) {  // the same for both lanes of the pair
    // singular face metric (polar axis): zero flux and wave speed
    if (side == 0) al[id] = 0.0;
#pragma unroll
    for (int q = 0; q < NC; ++q)
      if ((q < NH) == (side == 0)) glb(F.flux[DIR - 1][b * NC + q])[id] = 0.0;
    return;
  }
  const double gam = C.ph.gamma;
  St sm;
  calc_state(p, g, gc0, alpha, sm);
  double vmax, vmin;
  fast_ms<DIR>(gam, p, sm, gc0[0], gc0[DIR], gdd, vmax, vmin);
  const double vmaxO = pair_swap(vmax), vminO = pair_swap(vmin);
  const double vmaxL = side ? vmaxO : vmax, vminL = side ? vminO : vmin;
  const double vmaxR = side ? vmax : vmaxO, vminR = side ? vmin : vminO;
  const double cmax = fabs(kmax(kmax(0., vmaxL), vmaxR));
  const double cmin = fabs(kmax(kmax(0., -vminL), -vminR));
  const double cc = kmax(cmax, cmin);
  if (side == 0) al[id] = cc;
  double Um[NPRIM], Fm[NPRIM];
  contravariant_flux<0>(gam, p, sm, sg, Um);
  contravariant_flux<DIR>(gam, p, sm, sg, Fm);
  const double den = cmax + cmin + 1e-20, rden = 1.0 / den;
#pragma unroll
  for (int q = 0; q < NC; ++q) {
    const int s = slot_of<HB, HK>(q);
    const double Uo = pair_swap(Um[s]), Fo = pair_swap(Fm[s]);
    if ((q < NH) == (side == 0)) {
      const double UL = side ? Uo : Um[s], UR = side ? Um[s] : Uo;
      const double FL = side ? Fo : Fm[s], FR = side ? Fm[s] : Fo;
      double f;
      if (RS == PANGU_RIEMANN_LAXF)
        f = 0.5 * (FL + FR - cc * (UR - UL));
      else
        f = div_by(cmax * FL + cmin * FR - cmax * cmin * (UR - UL), den, rden);
      glb(F.flux[DIR - 1][b * NC + q])[id] = f;
    }
  }
}

// X1 sweep with one reconstruction per cell.  Lane l of a wavefront
// reconstructs cell c (both interface values, the same PPM4 / MC call the
// face kernel makes) and takes the right interface value of cell c - 1 from
// lane l - 1, so each cell's reconstruction runs once instead of twice.
// Waves cover 64 consecutive cells of a plane's row-major cell range
// (i = s - 1 .. e + 1) and overlap by one cell, so both neighbours of every
// face sit in one wave.  Same arithmetic as k_flux<1>: identical results.
template <int LIM, int RS, bool HB, bool HK>
__global__ __launch_bounds__(256) void k_flux_x1s(KCtx C, KFields F, double *const *alfven, Range3 R) {
  const int N = R.ntiles;
  const int n8 = (N + 7) >> 3;
  const int L = (int)blockIdx.x;
  const int T = (L & 7) * n8 + (L >> 3);
  if (T >= N) return;  // uniform over the workgroup
  const int b = R.tblock[T];
  const int local = T - R.tfirst[b];
  const DevBlock &B = C.blk[b];
  int lo[3], cnt[3];
  block_range(B, R, lo, cnt);
  const int plane = cnt[0] * cnt[1];
  const int wpp = (plane - 1 + 62) / 63;
  const int tpp = (wpp + 3) / 4;
  const int kk = local / tpp;
  const int w = (local - kk * tpp) * 4 + ((int)threadIdx.x >> 6);
  const int lane = (int)threadIdx.x & 63;
  const int cidx = w * 63 + lane;
  const bool have = (w < wpp) && (cidx < plane);
  const int jj = have ? cidx / cnt[0] : 0;
  const int ic = have ? cidx - jj * cnt[0] : 0;
  const int k = lo[2] + kk, j = lo[1] + jj, i = lo[0] + ic;
  const int64_t id = ((int64_t)k * B.n[1] + j) * B.n[0] + i;
  constexpr int NC = ncomp<HB, HK>();
  double rv[NPRIM], pl[NPRIM], pr[NPRIM];
#pragma unroll
  for (int s = 0; s < NPRIM; ++s) {
    rv[s] = 0.0;
    pl[s] = 0.0;
    pr[s] = 0.0;
  }
  // the cell stencils of every variable are loaded before the first
  // reconstruction (one wait for memory), then each cell's right interface
  // value moves one lane up through a DPP wavefront shift
  constexpr int NS = (LIM == PANGU_LIMITER_PPM4) ? 5 : 3;  // offsets -NS/2 .. NS/2
  double v[NC][NS];
#pragma unroll
  for (int q = 0; q < NC; ++q) {
    const auto a = glb(F.prim[b * NC + q]);
#pragma unroll
    for (int n = 0; n < NS; ++n) v[q][n] = have ? a[id + (n - NS / 2)] : 0.0;
  }
#pragma unroll
  for (int q = 0; q < NC; ++q) {
    const int s = slot_of<HB, HK>(q);
    if (LIM == PANGU_LIMITER_PPM4)
      recon_ppm4(v[q][0], v[q][1], v[q][2], v[q][3], v[q][4], rv[s], pr[s]);
    else
      recon_mc(v[q][0], v[q][1], v[q][2], rv[s], pr[s]);
    pl[s] = wave_shr1(rv[s]);
  }
  if (!have || lane == 0 || ic == 0) return;
  face_flux<1, RS, HB, HK>(C, F, alfven, b, k, j, i, pl, pr);
}

// Marching sweep along DIR (2 or 3): each thread walks R.step consecutive
// faces of one column (decode_march) and reconstructs each cell of the
// column once.  Cell m's reconstruction gives the right state of face m
// and the left state of face m + 1; the latter waits in LDS (one slot per
// thread, read back by the same thread) while face m's physics runs, so
// the registers peak as in k_flux.  A column of s faces reconstructs s + 1
// cells instead of 2s.  Same inputs and arithmetic per face as k_flux:
// identical results.  On an MI355X it ran 15-35 % slower than k_flux on
// 64x64x32 blocks: a cell's reconstruction is only ~13 % of a face's FP64
// instructions, and each face's loads wait for the previous face's physics
// instead of overlapping other faces' work; the autotuner picks it only
// where it wins.
template <int DIR, int LIM, int RS, bool HB, bool HK>
__global__ __launch_bounds__(256, 3) void k_flux_march(KCtx C, KFields F, double *const *alfven,
                                                    Range3 R) {
  int b = 0, k = 0, j = 0, i = 0, mend = 0;
  if (!decode_march<DIR>(C, R, b, k, j, i, mend)) return;
  const DevBlock &B = C.blk[b];
  const int64_t sj = B.n[0], sk = (int64_t)B.n[0] * B.n[1];
  const int64_t st = (DIR == 2) ? sj : sk;
  constexpr int NC = ncomp<HB, HK>();
  constexpr int NS = (LIM == PANGU_LIMITER_PPM4) ? 5 : 3;  // offsets -NS/2 .. NS/2
  __shared__ double carry[NPRIM][TILE];
  const int tid = (int)threadIdx.x;
  double pl[NPRIM], pr[NPRIM];
#pragma unroll
  for (int s = 0; s < NPRIM; ++s) {
    pl[s] = 0.0;
    pr[s] = 0.0;
  }
  int &m = (DIR == 2) ? j : k;
  // the left state of the first face: cell m - 1 at its right face
  {
    const int64_t id = (int64_t)k * sk + (int64_t)j * sj + i - st;
    double v[NC][NS];
#pragma unroll
    for (int q = 0; q < NC; ++q) {
      const auto a = glb(F.prim[b * NC + q]);
#pragma unroll
      for (int n = 0; n < NS; ++n) v[q][n] = a[id + (n - NS / 2) * st];
    }
#pragma unroll
    for (int q = 0; q < NC; ++q) {
      const int s = slot_of<HB, HK>(q);
      double t0;
      if (LIM == PANGU_LIMITER_PPM4)
        recon_ppm4(v[q][0], v[q][1], v[q][2], v[q][3], v[q][4], pl[s], t0);
      else
        recon_mc(v[q][0], v[q][1], v[q][2], pl[s], t0);
    }
  }
#pragma unroll 1
  for (; m < mend; ++m) {
    // an opaque step: every address below is formed from m afresh, so the
    // compiler keeps no per-array pointer alive across the iterations
    __asm__ volatile("" : "+v"(m));
    const int64_t id = (int64_t)k * sk + (int64_t)j * sj + i;
    double v[NC][NS];
#pragma unroll
    for (int q = 0; q < NC; ++q) {
      const auto a = glb(F.prim[b * NC + q]);
#pragma unroll
      for (int n = 0; n < NS; ++n) v[q][n] = a[id + (n - NS / 2) * st];
    }
#pragma unroll
    for (int q = 0; q < NC; ++q) {
      const int s = slot_of<HB, HK>(q);
      double nx;
      if (LIM == PANGU_LIMITER_PPM4)
        recon_ppm4(v[q][0], v[q][1], v[q][2], v[q][3], v[q][4], nx, pr[s]);
      else
        recon_mc(v[q][0], v[q][1], v[q][2], nx, pr[s]);
      carry[s][tid] = nx;
    }
    // keep the carried states in LDS, not in registers, across the face
    // physics (the compiler would otherwise forward the stores)
    __asm__ volatile("" ::: "memory");
    face_flux<DIR, RS, HB, HK>(C, F, alfven, b, k, j, i, pl, pr);
    __asm__ volatile("" ::: "memory");
#pragma unroll
    for (int q = 0; q < NC; ++q) {
      const int s = slot_of<HB, HK>(q);
      pl[s] = carry[s][tid];
    }
  }
}

// ------------------------------------------------------ constrained transport
// Corner EMFs (constrained_transport.cc:49-76).
__global__ __launch_bounds__(256) void k_emf(KCtx C, KFields F, double *const *emf, Range3 R) {
  int b, k, j, i;
  if (!decode_tile(C, R, b, k, j, i)) return;
  const DevBlock &B = C.blk[b];
  const bool o1 = B.n[0] > 1, o2 = B.n[1] > 1, o3 = B.n[2] > 1;
  const bool e1on = o2 && o3, e2on = o3 && o1, e3on = o1 && o2;
  const int64_t sj = B.n[0], sk = (int64_t)B.n[0] * B.n[1];
  const int64_t id = k * sk + j * sj + i;
  const int nc = C.nc;
  const int b1 = C.iBX, b2 = C.iBX + 1, b3 = C.iBX + 2;
  double e1 = 0.0, e2 = 0.0, e3 = 0.0;
  if (e1on) {
    const auto f2 = glb(F.flux[1][b * nc + b3]), f3 = glb(F.flux[2][b * nc + b2]);
    e1 = 0.25 * ((f2[id] + f2[id - sk]) - (f3[id] + f3[id - sj]));
  }
  if (e2on) {
    const auto f3 = glb(F.flux[2][b * nc + b1]), f1 = glb(F.flux[0][b * nc + b3]);
    e2 = 0.25 * ((f3[id] + f3[id - 1]) - (f1[id] + f1[id - sk]));
  }
  if (e3on) {
    const auto f1 = glb(F.flux[0][b * nc + b2]), f2 = glb(F.flux[1][b * nc + b1]);
    e3 = 0.25 * ((f1[id] + f1[id - sj]) - (f2[id] + f2[id - 1]));
  }
  glb(emf[b * 3 + 0])[id] = e1;
  glb(emf[b * 3 + 1])[id] = e2;
  glb(emf[b * 3 + 2])[id] = e3;
}

// Face B-flux overwrite from the EMFs (constrained_transport.cc:78-127).
// One launch covers the union of the three face ranges.
__global__ __launch_bounds__(256) void k_ct_faces(KCtx C, KFields F, double *const *emf, Range3 R) {
  int b, k, j, i;
  if (!decode_tile(C, R, b, k, j, i)) return;
  const DevBlock &B = C.blk[b];
  const int o1 = B.n[0] > 1, o2 = B.n[1] > 1, o3 = B.n[2] > 1;
  const int64_t sj = B.n[0], sk = (int64_t)B.n[0] * B.n[1];
  const int64_t id = k * sk + j * sj + i;
  const int nc = C.nc;
  const int b1 = C.iBX, b2 = C.iBX + 1, b3 = C.iBX + 2;
  const int E1 = o2 && o3, E2 = o3 && o1, E3 = o1 && o2;
  const auto e1 = glb(emf[b * 3 + 0]), e2 = glb(emf[b * 3 + 1]), e3 = glb(emf[b * 3 + 2]);
  const bool in_k = k <= B.e[2], in_j = j <= B.e[1], in_i = i <= B.e[0];
  if (in_k && in_j) {  // X1 faces: i in [is, ie+o1]
    if (o1) glb(F.flux[0][b * nc + b1])[id] = 0;
    if (E2) glb(F.flux[0][b * nc + b3])[id] = -0.5 * (e2[id] + e2[id + sk]);
    if (E3) glb(F.flux[0][b * nc + b2])[id] = 0.5 * (e3[id] + e3[id + sj]);
  }
  if (in_k && in_i) {  // X2 faces: j in [js, je+o2]
    if (E1) glb(F.flux[1][b * nc + b3])[id] = 0.5 * (e1[id] + e1[id + sk]);
    if (o2) glb(F.flux[1][b * nc + b2])[id] = 0.;
    if (E3) glb(F.flux[1][b * nc + b1])[id] = -0.5 * (e3[id] + e3[id + 1]);
  }
  if (in_j && in_i && o3) {  // X3 faces: k in [ks, ke+o3]
    if (E1) glb(F.flux[2][b * nc + b2])[id] = -0.5 * (e1[id] + e1[id + sj]);
    if (E2) glb(F.flux[2][b * nc + b1])[id] = 0.5 * (e2[id] + e2[id + 1]);
    glb(F.flux[2][b * nc + b3])[id] = 0;
  }
}

// ------------------------------------ flux divergence, update, source term
// dU/dt = -[sum_d (A_d F_d(+) - A_d F_d(-))]/V (Parthenon FluxDivergence),
// U(out) = U(init) + beta dt dU/dt (UpdateIndependentData) and the
// geometric source U_nu += beta dt sqrt|g| T^r_c Gamma^c_{nu r}
// (source_term.cc:171-228).  MODE: 1 = divergence only (writes dU/dt),
// 2 = update only (reads dU/dt), 3 = source only, 4 = all three fused.
// Flux divergence of one cell (and, for MODE 4, the update from the initial
// container) with the X2 / X3 terms present when D2 / D3.
// CTF: the face B fluxes are the constrained-transport values, formed here
// from edge EMFs that are themselves formed here from the Riemann B fluxes,
// exactly as k_emf and k_ct_faces would compute them (same operands, same
// operations), so neither CT kernel runs.  Each cell forms the 12 EMFs on
// its edges (each edge is shared by four cells; the extra arithmetic is
// cheaper than a kernel that writes and rereads the EMF arrays).  B
// components sit at packed indices 5..7.
template <bool D2, bool D3, int MODE, int NC, bool CTF>
__device__ __forceinline__ void flux_div(const DevBlock &B, const KFields &F0, const KFields &Finit,
                                         double *const *dudt, int b,
                                         int64_t id, int64_t sj, int64_t sk, double bdt,
                                         double unew[NC]) {
  const double a1 = B.dx[1] * B.dx[2], a2 = B.dx[0] * B.dx[2], a3 = B.dx[0] * B.dx[1];
  const double vol = B.dx[0] * B.dx[1] * B.dx[2];
  const double rvol = 1.0 / vol;
  double f1p[NC], f1m[NC], f2p[NC], f2m[NC], f3p[NC], f3m[NC], u0[NC];
  // Every access is base[u + id]: u is a workgroup-uniform offset (a row
  // sj, a plane sk, or 1) folded into the scalar base, id the cell.
  const uint32_t id32 = (uint32_t)id;
  auto at = [&](const double *p, int64_t u) { return ld32(p, u, id32); };
  // edge EMFs at uniform offset u from the cell, with the k_emf expressions
  const double *f1b2 = nullptr, *f1b3 = nullptr, *f2b1 = nullptr, *f2b3 = nullptr, *f3b1 = nullptr,
               *f3b2 = nullptr;
  if (CTF) {
    f1b2 = F0.flux[0][b * NC + 6];
    f1b3 = F0.flux[0][b * NC + 7];
    if (D2) {
      f2b1 = F0.flux[1][b * NC + 5];
      f2b3 = F0.flux[1][b * NC + 7];
    }
    if (D3) {
      f3b1 = F0.flux[2][b * NC + 5];
      f3b2 = F0.flux[2][b * NC + 6];
    }
  }
  auto E1 = [&](int64_t u) { return 0.25 * ((at(f2b3, u) + at(f2b3, u - sk)) - (at(f3b2, u) + at(f3b2, u - sj))); };
  auto E2 = [&](int64_t u) { return 0.25 * ((at(f3b1, u) + at(f3b1, u - 1)) - (at(f1b3, u) + at(f1b3, u - sk))); };
  auto E3 = [&](int64_t u) { return 0.25 * ((at(f1b2, u) + at(f1b2, u - sj)) - (at(f2b1, u) + at(f2b1, u - 1))); };
#pragma unroll
  for (int q = 0; q < NC; ++q) {
    const int c = q - 5;  // B component when CTF and 0 <= c < 3
    const bool ctb = CTF && c >= 0 && c < 3;
    const double *f1 = F0.flux[0][b * NC + q];
    if (ctb && c == 0) {  // normal component: zero
      f1p[q] = 0;
      f1m[q] = 0;
    } else if (ctb && c == 1 && D2) {
      f1m[q] = 0.5 * (E3(0) + E3(sj));
      f1p[q] = 0.5 * (E3(1) + E3(1 + sj));
    } else if (ctb && c == 2 && D3) {
      f1m[q] = -0.5 * (E2(0) + E2(sk));
      f1p[q] = -0.5 * (E2(1) + E2(1 + sk));
    } else {
      f1p[q] = glb(f1)[id + 1];
      f1m[q] = glb(f1)[id];
    }
    if (D2) {
      const double *f2 = F0.flux[1][b * NC + q];
      if (ctb && c == 0) {
        f2m[q] = -0.5 * (E3(0) + E3(1));
        f2p[q] = -0.5 * (E3(sj) + E3(sj + 1));
      } else if (ctb && c == 1) {
        f2p[q] = 0.;
        f2m[q] = 0.;
      } else if (ctb && c == 2 && D3) {
        f2m[q] = 0.5 * (E1(0) + E1(sk));
        f2p[q] = 0.5 * (E1(sj) + E1(sj + sk));
      } else {
        f2p[q] = at(f2, sj);
        f2m[q] = at(f2, 0);
      }
    }
    if (D3) {
      const double *f3 = F0.flux[2][b * NC + q];
      if (ctb && c == 0) {
        f3m[q] = 0.5 * (E2(0) + E2(1));
        f3p[q] = 0.5 * (E2(sk) + E2(sk + 1));
      } else if (ctb && c == 1 && D2) {
        f3m[q] = -0.5 * (E1(0) + E1(sj));
        f3p[q] = -0.5 * (E1(sk) + E1(sk + sj));
      } else if (ctb && c == 2) {
        f3p[q] = 0;
        f3m[q] = 0;
      } else {
        f3p[q] = at(f3, sk);
        f3m[q] = at(f3, 0);
      }
    }
    if (MODE == 4) u0[q] = at(Finit.cons[b * NC + q], 0);
  }
#pragma unroll
  for (int q = 0; q < NC; ++q) {
    double s = (a1 * f1p[q] - a1 * f1m[q]);
    if (D2) s += (a2 * f2p[q] - a2 * f2m[q]);
    if (D3) s += (a3 * f3p[q] - a3 * f3m[q]);
    const double du = div_by(-s, vol, rvol);
    if (MODE == 1)
      glb(dudt[b * NC + q])[id] = du;
    else
      unew[q] = u0[q] + bdt * du;
  }
}

// Five waves per SIMD is the most the magnetised source-term variants reach
// without spilling (the field-free ones spill at five and run at four); the
// kernel is memory bound, so occupancy pays.
// One cell of k_div_update; the new U also goes to uout (when given).
template <int MODE, bool SRC, bool HB, bool HK, bool CTF>
__device__ __forceinline__ void div_update_cell(const KCtx &C, const KFields &F0, const KFields &Finit,
                                                const KFields &F1, double *const *dudt,
                                                const double *dt_dev, double beta, double bdt_host,
                                                int b, int k, int j, int i, double *uout) {
  const DevBlock &B = C.blk[b];
  const int64_t sj = B.n[0], sk = (int64_t)B.n[0] * B.n[1];
  const int64_t id = k * sk + j * sj + i;
  constexpr int NC = ncomp<HB, HK>();
  const double bdt = dt_dev ? beta * dt_dev[2 * B.mesh] : bdt_host;
  double unew[NC];
  if (MODE == 1 || MODE == 4) {
    // one uniform branch on the block's active dimensions, so the loads of
    // all components sit in one basic block and overlap
    const bool d2 = B.n[1] > 1, d3 = B.n[2] > 1;
    if (d2 && d3)
      flux_div<true, true, MODE, NC, CTF>(B, F0, Finit, dudt, b, id, sj, sk, bdt, unew);
    else if (d2)
      flux_div<true, false, MODE, NC, CTF>(B, F0, Finit, dudt, b, id, sj, sk, bdt, unew);
    else if (d3)
      flux_div<false, true, MODE, NC, CTF>(B, F0, Finit, dudt, b, id, sj, sk, bdt, unew);
    else
      flux_div<false, false, MODE, NC, CTF>(B, F0, Finit, dudt, b, id, sj, sk, bdt, unew);
  } else if (MODE == 2) {
#pragma unroll
    for (int q = 0; q < NC; ++q)
      unew[q] = glb(Finit.cons[b * NC + q])[id] + bdt * glb(dudt[b * NC + q])[id];
  } else {
#pragma unroll
    for (int q = 0; q < NC; ++q) unew[q] = glb(F1.cons[b * NC + q])[id];
  }
  if (MODE == 1) return;
  if (SRC && (MODE == 3 || MODE == 4)) {
    const DevGeo &G = C.geo[B.geo];
    const int64_t t = cen_index(G, k, j, i);
    CenLoad m;
    load_cen_state(G, t, m);
    double p[NPRIM];
#pragma unroll
    for (int s = 0; s < NPRIM; ++s) p[s] = 0.0;
#pragma unroll
    for (int q = 0; q < NC; ++q) p[slot_of<HB, HK>(q)] = glb(F1.prim[b * NC + q])[id];
    St s;
    calc_state(p, m.g, m.gc0, m.alpha, s);
    const double gam = C.ph.gamma;
    const double pg = (gam - 1.) * p[ENY];
    const double h = p[RHO] + gam * p[ENY];
    const double etot = h + s.bsq, ptot = pg + 0.5 * s.bsq;
    double T[4][4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      T[0][c] = tmunu<0>(s, etot, ptot, c);
      T[1][c] = tmunu<1>(s, etot, ptot, c);
      T[2][c] = tmunu<2>(s, etot, ptot, c);
      T[3][c] = tmunu<3>(s, etot, ptot, c);
    }
    // one scalar base for the connection block (40 or 64 slots) and a
    // 32-bit index per slot: per-slot scalar bases would exhaust the scalar
    // registers
    const double *conn = G.cen + (int64_t)CG_CONN * G.ncen;
    const uint32_t ncen32 = (uint32_t)G.ncen, t32 = (uint32_t)t;
    const bool sym = G.conn_sym != 0;
#pragma unroll
    for (int dd = 0; dd < 4; ++dd) {
      double con = 0.0;
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const int slot = sym ? 10 * c + conn_pair(dd, r) : 16 * c + 4 * dd + r;
          con += T[r][c] * ldc32(conn, 0, (uint32_t)slot * ncen32 + t32);
        }
      const int q = (dd == 0) ? 1 : 1 + dd;
      unew[q] += con * m.sg * bdt;
    }
  }
#pragma unroll
  for (int q = 0; q < NC; ++q) glb(F1.cons[b * NC + q])[id] = unew[q];
  if (uout) {
#pragma unroll
    for (int q = 0; q < NC; ++q) uout[q] = unew[q];
  }
}

template <int MODE, bool SRC, bool HB, bool HK, bool CTF>
__global__ __launch_bounds__(256, (HB && MODE != 1) ? 5 : 4) void k_div_update(KCtx C, KFields F0, KFields Finit, KFields F1,
                                                    double *const *dudt,
                                                    const double *dt_dev, double beta,
                                                    double bdt_host, Range3 R) {
  int b, k, j, i;
  if (!decode_tile(C, R, b, k, j, i)) return;
  div_update_cell<MODE, SRC, HB, HK, CTF>(C, F0, Finit, F1, dudt, dt_dev, beta, bdt_host, b, k, j, i,
                                          nullptr);
}

// -------------------------------------------------- recovery (+ heating)
// ApplyElectronHeating (electron_heating.cc:66-109) on one recovered cell
// from its primitives p (rho, u, u~, B and the stale entropy; K_e is not
// read); the centre metric is loaded here.
template <bool HB, bool HK>
__device__ __forceinline__ void heat_vals(const KCtx &C, const KFields &F, const DevGeo &G, int64_t t,
                                          int b, int64_t id, const Arr32<double> *pp,
                                          const double p[NPRIM], const double *ureg = nullptr) {
  constexpr int NC = ncomp<HB, HK>();
  CenLoad m;
  load_cen_state(G, t, m);
  St s;
  calc_state(p, m.g, m.gc0, m.alpha, s);
  const uint32_t id32 = (uint32_t)id;
  // conserved components: from memory, or from registers (ureg, in
  // component order) when the caller has just written them
  auto cu = [&](int q) { return ureg ? ureg[q] : ld32(F.cons[b * NC + q], 0, id32); };
  const double crho = cu(0);
  HeatIn c;
  c.rho = p[RHO];
  c.energy = p[ENY];
  c.bsq = s.bsq;
  c.adv_total = cu(ient<HB>()) / crho;
  c.rec_total = p[ENT];
  c.adv_e = cu(ikel<HB>()) / crho;
  pp[ikel<HB>()][id] = heat_apply(C.ph, c);
}

// The same with the primitives read from memory.
template <bool HB, bool HK>
__device__ __forceinline__ void heat_cell(const KCtx &C, const KFields &F, const DevGeo &G, int64_t t,
                                          int b, int64_t id, const Arr32<double> *pp) {
  constexpr int NC = ncomp<HB, HK>();
  double p[NPRIM];
#pragma unroll
  for (int s = 0; s < NPRIM; ++s) p[s] = 0.0;
#pragma unroll
  for (int q = 0; q < NC; ++q) p[slot_of<HB, HK>(q)] = pp[q][id];
  heat_vals<HB, HK>(C, F, G, t, b, id, pp, p);
}

// Recovery (invertor.cc:184-319) and, when HEAT, ApplyElectronHeating
// (electron_heating.cc:66-109) on the recovered cell.  REC: 0 chainer,
// 1 Kastaun, 2 1Dvsq alone, -1 none (heating only).
template <int REC, bool HB, bool HK, bool HEAT>
__device__ __forceinline__ void recover_cell(const KCtx &C, const KFields &F, const SharedTabs &S, int b,
                                             int k, int j, int i, double x0 = -1.0,
                                             const double *ureg = nullptr) {
  const DevBlock &B = C.blk[b];
  const int64_t id = ((int64_t)k * B.n[1] + j) * B.n[0] + i;
  constexpr int NC = ncomp<HB, HK>();
  const DevGeo &G = C.geo[B.geo];
  const int64_t t = cen_index(G, k, j, i);
  const int64_t ns = G.ncen;
  const Phys &P = C.ph;
  Arr32<double> pp[NC];
#pragma unroll
  for (int q = 0; q < NC; ++q) pp[q] = Arr32<double>{F.prim[b * NC + q]};
  if (REC >= 0) {
    const bool chainer = (REC != 1);
    bool excised = false;
    if (P.metric == PANGU_METRIC_CKS && ldc32(G.cen, (int64_t)CG_RAD * ns, (uint32_t)t) < P.r_excise) excised = true;
    if (excised) {
      if (chainer && S.flag) glb(S.flag[b])[id] = 0.0;
      if (S.status) glb(S.status[b])[id] = -1;
      if (S.iters) glb(S.iters[b])[id] = 0;
    } else {
      CenMetric m;
#pragma unroll
      for (int n = 0; n < 16; ++n) {
        m.gcov[n] = ldc32(G.cen, (int64_t)(CG_GCOV + n) * ns, (uint32_t)t);
        m.gcon[n] = ldc32(G.cen, (int64_t)(CG_GCON + n) * ns, (uint32_t)t);
      }
      m.lapse = ldc32(G.cen, (int64_t)CG_LAPSE * ns, (uint32_t)t);
      const double sg = ldc32(G.cen, (int64_t)CG_SG * ns, (uint32_t)t);
      const double alpha = ldc32(G.cen, (int64_t)CG_ALPHA * ns, (uint32_t)t);
      double U[9], Ph[9], bnew[3] = {0.0, 0.0, 0.0};
      Arr32<double> cons[NC];
#pragma unroll
      for (int q = 0; q < NC; ++q) cons[q] = Arr32<double>{F.cons[b * NC + q]};
      // U from memory, or from registers (ureg) when the update just wrote it
      auto cv = [&](int q) { return ureg ? ureg[q] : (double)cons[q][id]; };
      const double c_rho = cv(0), c_eny = cv(1);
      const double isg = 1.0 / sg;
      const double aos = alpha * isg;
      U[RHO] = aos * c_rho;
      U[ENY] = aos * (c_eny - c_rho);
      U[UX1] = aos * cv(2);
      U[UX2] = aos * cv(3);
      U[UX3] = aos * cv(4);
      if (HB) {
        const double cb1 = cv(5), cb2 = cv(6), cb3 = cv(7);
        U[BX1] = aos * cb1;
        U[BX2] = aos * cb2;
        U[BX3] = aos * cb3;
        bnew[0] = cb1 * isg;
        bnew[1] = cb2 * isg;
        bnew[2] = cb3 * isg;
        pp[5][id] = bnew[0];
        pp[6][id] = bnew[1];
        pp[7][id] = bnew[2];
      } else {
        U[BX1] = 0.0;
        U[BX2] = 0.0;
        U[BX3] = 0.0;
      }
      U[ENT] = aos * cv(ient<HB>());
      // the Newton schemes start from the previous primitives; Kastaun's
      // bracket needs none (it writes rho, u and u~ only on success), so
      // their loads are skipped
#pragma unroll
      for (int s = 0; s < 9; ++s) Ph[s] = 0.0;
      if (REC != 1) {
        Ph[RHO] = pp[0][id];
        Ph[ENY] = pp[1][id];
        Ph[UX1] = pp[2][id];
        Ph[UX2] = pp[3][id];
        Ph[UX3] = pp[4][id];
        Ph[ENT] = pp[ient<HB>()][id];
      }
      int st = 0, it = 0, itn = 0;
      bool write = true;
      if (REC == 1) {
        double mu = 0.0;
        st = c2p_kastaun(U, Ph, P.gamma, m, it, &mu);
        if (S.root) glb(S.root[b])[id] = (st == 0) ? mu : __builtin_nan("");
        itn = it;
        write = (st == 0);
      } else if (REC == 2) {
        // x0 >= 0: the learned start (k_recover_learned_vsq)
        st = c2p_1dvsq(U, Ph, P.gamma, m, it, x0);
        itn = it;
      } else {
        st = c2p_2d(U, Ph, P.gamma, m, it);
        itn = it;
        if (st != 0) {
          st = c2p_1dvsq(U, Ph, P.gamma, m, it);
          itn += it;
          if (st != 0) {
            st = c2p_1d(U, Ph, P.gamma, m, it);
            itn += it;
          }
        }
      }
      if (chainer && S.flag) {
        const double fv = (st == 0) ? 1.0 : 0.0;
        glb(S.flag[b])[id] = fv;
        if (S.flag_rec) glb(S.flag_rec[b])[id] = fv;
      }
      if (S.status) glb(S.status[b])[id] = st;
      if (S.iters) glb(S.iters[b])[id] = itn;
      if (write) {
        pp[0][id] = Ph[RHO];
        pp[1][id] = Ph[ENY];
        pp[2][id] = Ph[UX1];
        pp[3][id] = Ph[UX2];
        pp[4][id] = Ph[UX3];
      }
      if (HEAT && HK) {
        // the heating reads the primitives just written from registers (the
        // compiler cannot forward them: the stores may alias other arrays);
        // a cell whose recovery failed keeps its previous primitives
        double hp[NPRIM];
#pragma unroll
        for (int s = 0; s < NPRIM; ++s) hp[s] = 0.0;
        if (write) {
          hp[RHO] = Ph[RHO];
          hp[ENY] = Ph[ENY];
          hp[UX1] = Ph[UX1];
          hp[UX2] = Ph[UX2];
          hp[UX3] = Ph[UX3];
        } else {
          hp[RHO] = pp[0][id];
          hp[ENY] = pp[1][id];
          hp[UX1] = pp[2][id];
          hp[UX2] = pp[3][id];
          hp[UX3] = pp[4][id];
        }
        if (HB) {
          hp[BX1] = bnew[0];
          hp[BX2] = bnew[1];
          hp[BX3] = bnew[2];
        }
        hp[ENT] = pp[ient<HB>()][id];
        heat_vals<HB, HK>(C, F, G, t, b, id, pp, hp, ureg);
        return;
      }
    }
  }
  if (HEAT && HK) heat_cell<HB, HK>(C, F, G, t, b, id, pp);
}

template <int REC, bool HB, bool HK, bool HEAT>
__global__ __launch_bounds__(256) void k_recover(KCtx C, KFields F, SharedTabs S, Range3 R) {
  int b, k, j, i;
  if (!decode_tile(C, R, b, k, j, i)) return;
  recover_cell<REC, HB, HK, HEAT>(C, F, S, b, k, j, i);
}

// The update of k_div_update (mode 4: flux divergence, RK update, source)
// and the recovery (+ heating) of k_recover for the same cell in one pass:
// the recovery takes the new U from registers instead of memory.  The update
// is memory-bound and the recovery VALU-bound, so their phases overlap
// across the wavefronts of one launch.  Same operations per cell in the same
// order: identical results.  Called only when nothing has to run between
// the two (the stage's exchange runs before).
template <bool SRC, bool HB, bool HK, bool CTF, int REC, bool HEAT>
__global__ __launch_bounds__(256) void k_update_recover(KCtx C, KFields F0, KFields Finit, KFields F1,
                                                        SharedTabs S, const double *dt_dev, double beta,
                                                        double bdt_host, Range3 R) {
  int b, k, j, i;
  if (!decode_tile(C, R, b, k, j, i)) return;
  double u[NPRIM];
  div_update_cell<4, SRC, HB, HK, CTF>(C, F0, Finit, F1, nullptr, dt_dev, beta, bdt_host, b, k, j, i, u);
  recover_cell<REC, HB, HK, HEAT>(C, F1, S, b, k, j, i, -1.0, u);
}

// Persistent form of k_recover: a grid no larger than what the GPU holds
// at once pulls tiles from a queue (queue[0]: next tile, queue[1]:
// workgroups done) until every tile is taken; the last workgroup out resets
// the queue for the next launch on the stream, so a captured graph replays
// it without a memset.  Tiles carry no dependencies, so no workgroup ever
// waits for another.  Same cells, same arithmetic: identical results.
template <int REC, bool HB, bool HK, bool HEAT>
__global__ __launch_bounds__(256) void k_recover_queue(KCtx C, KFields F, SharedTabs S, Range3 R,
                                                       unsigned *queue) {
  __shared__ int tile_sh;
  for (;;) {
    if (threadIdx.x == 0) tile_sh = (int)atomicAdd(&queue[0], 1u);
    __syncthreads();
    const int T = tile_sh;
    __syncthreads();  // every thread has read the tile before the next pull
    if (T >= R.ntiles) break;
    int b, k, j, i;
    if (decode_tile_at(C, R, T, b, k, j, i)) recover_cell<REC, HB, HK, HEAT>(C, F, S, b, k, j, i);
  }
  if (threadIdx.x == 0 && atomicAdd(&queue[1], 1u) == gridDim.x - 1) {
    atomicExch(&queue[0], 0u);
    atomicExch(&queue[1], 0u);
  }
}

// Kastaun recovery with the learned starting bracket (and, when HEAT, the
// electron heating): k_recover<1, ...> whose wavefronts that hold a cell
// passing the gate (its previous recovery took at least S.gate
// false-position steps) run the network on the matrix cores inside the
// wavefront (wave_mlp); every cell of such a wavefront then starts the
// search from the prediction (kastaun_guess_k), the others search [0, 1]
// as k_recover does.  Status, iteration and field outputs are those of
// k_recover; the recovered state agrees with it to the solver tolerance
// (another false-position path to a root of the same function), not bit for
// bit.
template <bool HB, bool HK>
__device__ __forceinline__ int learned_setup(const KCtx &C, const KFields &F, const DevGeo &G, int64_t t,
                                             int b, int64_t id, KastaunIn &kin) {
  constexpr int NC = ncomp<HB, HK>();
  const int64_t ns = G.ncen;
  CenMetric m;
#pragma unroll
  for (int n = 0; n < 16; ++n) {
    m.gcov[n] = ldc32(G.cen, (int64_t)(CG_GCOV + n) * ns, (uint32_t)t);
    m.gcon[n] = ldc32(G.cen, (int64_t)(CG_GCON + n) * ns, (uint32_t)t);
  }
  m.lapse = ldc32(G.cen, (int64_t)CG_LAPSE * ns, (uint32_t)t);
  const double aos = ldc32(G.cen, (int64_t)CG_ALPHA * ns, (uint32_t)t) * (1.0 / ldc32(G.cen, (int64_t)CG_SG * ns, (uint32_t)t));
  const double *const *cons = F.cons + (int64_t)b * NC;
  const double c_rho = glb(cons[0])[id], c_eny = glb(cons[1])[id];
  double U[9];
  U[RHO] = aos * c_rho;
  U[ENY] = aos * (c_eny - c_rho);
  U[UX1] = aos * glb(cons[2])[id];
  U[UX2] = aos * glb(cons[3])[id];
  U[UX3] = aos * glb(cons[4])[id];
#pragma unroll
  for (int a = 0; a < 3; ++a) U[BX1 + a] = HB ? aos * glb(cons[5 + a])[id] : 0.0;
  U[ENT] = aos * glb(cons[ient<HB>()])[id];
  return kastaun_setup(U, m, kin);
}

// UPD 1 or 2: the cell's update (k_div_update mode 4; 2 with the source)
// runs first, as in k_update_recover; the recovery then reads the U this
// thread has just stored.
template <bool HB, bool HK, bool HEAT, int UPD>
__global__ __launch_bounds__(256) void k_recover_learned(KCtx C, KFields F, SharedTabs S, MlpFrag W,
                                                         Range3 R, UpdArgs A) {
  int b = 0, k = 0, j = 0, i = 0;
  const bool cell = decode_tile(C, R, b, k, j, i);  // false on the whole workgroup past the range
  if (UPD && cell)
    div_update_cell<4, UPD == 2, HB, HK, HB>(C, A.F0, A.Finit, F, nullptr, A.dt_dev, A.beta, A.bdt_host, b,
                                             k, j, i, nullptr);
  const int lane = threadIdx.x & 63;
  constexpr int NC = ncomp<HB, HK>();
  const Phys &P = C.ph;
  const DevBlock &B = C.blk[cell ? b : 0];
  const DevGeo &G = C.geo[B.geo];
  const int64_t id = cell ? ((int64_t)k * B.n[1] + j) * B.n[0] + i : 0;
  const int64_t t = cell ? cen_index(G, k, j, i) : 0;
  const bool active =
      cell && !(P.metric == PANGU_METRIC_CKS && ldc32(G.cen, (int64_t)CG_RAD * G.ncen, (uint32_t)t) < P.r_excise);
  // The workgroup votes on the previous iteration counts alone, before any
  // other load: a workgroup with no cell passing the gate then runs as
  // k_recover<1> does, with no further barrier.  The network's cost is per
  // wavefront, so a wavefront that runs it for a gated cell serves every
  // cell of it with a valid setup.
  int prev = 0;
  if (active && S.iters) prev = glb(S.iters[b])[id];
  const bool gated = active && S.iters && prev >= S.gate;
  KastaunIn kin;
  int setup = -1;
  double guess = -1.0;
  __shared__ MlpLds lds;
  // every thread of the workgroup reaches the vote (no early return above)
  if (__syncthreads_or(gated)) {
    mlp_stage(W, lds, threadIdx.x, blockDim.x);
    if (active) setup = learned_setup<HB, HK>(C, F, G, t, b, id, kin);
    __syncthreads();
    const bool want = setup == 0;
    if (__ballot(gated && want)) {  // uniform over the wavefront: all 64 lanes run the network
      float x[5] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
      if (want) {
        float f[5];
        c2p_features(kin, f);
#pragma unroll
        for (int q = 0; q < 5; ++q) x[q] = bf2f(f2bf((f[q] - W.mean[q]) / W.dev[q]));
      }
      const float z = wave_mlp(lds, W.b4[0], x, lane);
      if (want) guess = (double)(1.0f / (1.0f + expf(-z)));
      // the setup is recomputed from memory (same inputs, same bits)
      // instead of being held in registers across the network
      __asm__ volatile("" ::: "memory");
      if (active) setup = learned_setup<HB, HK>(C, F, G, t, b, id, kin);
    }
  } else if (active) {
    setup = learned_setup<HB, HK>(C, F, G, t, b, id, kin);
  }
  if (!cell) return;
  Arr32<double> pp[NC];
#pragma unroll
  for (int q = 0; q < NC; ++q) pp[q] = Arr32<double>{F.prim[b * NC + q]};
  if (!active) {
    if (S.status) glb(S.status[b])[id] = -1;
    if (S.iters) glb(S.iters[b])[id] = 0;
  } else {
    double bnew[3] = {0.0, 0.0, 0.0};
    if (HB) {
      const double isg = 1.0 / ldc32(G.cen, (int64_t)CG_SG * G.ncen, (uint32_t)t);
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        bnew[a] = glb(F.cons[b * NC + 5 + a])[id] * isg;
        pp[5 + a][id] = bnew[a];
      }
    }
    double Ph[9];
    double mu = 0.0;
    int st, it = 0;
    if (setup != 0) {
      st = setup;
    } else {
      // one code path: without a guess (-1) kastaun_guess_k searches [0, 1]
      // exactly as kastaun_plain_k does
      int ev = 0;
      st = kastaun_guess_k(kin, P.gamma, Ph, guess, S.delta, it, ev, &mu);
    }
    if (S.root) glb(S.root[b])[id] = (st == 0) ? mu : __builtin_nan("");
    if (S.status) glb(S.status[b])[id] = st;
    if (S.iters) glb(S.iters[b])[id] = it;
    if (st == 0) {
      pp[0][id] = Ph[RHO];
      pp[1][id] = Ph[ENY];
      pp[2][id] = Ph[UX1];
      pp[3][id] = Ph[UX2];
      pp[4][id] = Ph[UX3];
    }
    if (HEAT && HK) {
      // as recover_cell: the heating takes the primitives just written from
      // registers (a failed cell keeps its previous ones, from memory)
      double hp[NPRIM];
#pragma unroll
      for (int q = 0; q < NPRIM; ++q) hp[q] = 0.0;
      if (st == 0) {
        hp[RHO] = Ph[RHO];
        hp[ENY] = Ph[ENY];
        hp[UX1] = Ph[UX1];
        hp[UX2] = Ph[UX2];
        hp[UX3] = Ph[UX3];
      } else {
        hp[RHO] = pp[0][id];
        hp[ENY] = pp[1][id];
        hp[UX1] = pp[2][id];
        hp[UX2] = pp[3][id];
        hp[UX3] = pp[4][id];
      }
      if (HB) {
#pragma unroll
        for (int a = 0; a < 3; ++a) hp[BX1 + a] = bnew[a];
      }
      hp[ENT] = pp[ient<HB>()][id];
      heat_vals<HB, HK>(C, F, G, t, b, id, pp, hp);
      return;
    }
  }
  if (HEAT && HK) heat_cell<HB, HK>(C, F, G, t, b, id, pp);
}

// The 1Dvsq recovery with the learned start (and, when HEAT, the heating):
// the vote, the staging and the in-wave network of k_recover_learned, on
// the same features (the Kastaun invariants of the cell's U and metric),
// predicting z = log u^2; every cell of a wavefront that ran the network
// then starts its Newton iteration from v^2 = u^2 / (1 + u^2) =
// logistic(z) instead of the previous primitives' v^2.  Everything else is
// recover_cell<2>: the result agrees with k_recover<2> to the Newton
// tolerance, not bit for bit.
template <bool HB, bool HK, bool HEAT, int UPD>
__global__ __launch_bounds__(256) void k_recover_learned_vsq(KCtx C, KFields F, SharedTabs S, MlpFrag W,
                                                             Range3 R, UpdArgs A) {
  int b = 0, k = 0, j = 0, i = 0;
  const bool cell = decode_tile(C, R, b, k, j, i);  // false on the whole workgroup past the range
  if (UPD && cell)
    div_update_cell<4, UPD == 2, HB, HK, HB>(C, A.F0, A.Finit, F, nullptr, A.dt_dev, A.beta, A.bdt_host, b,
                                             k, j, i, nullptr);
  const int lane = threadIdx.x & 63;
  const Phys &P = C.ph;
  const DevBlock &B = C.blk[cell ? b : 0];
  const DevGeo &G = C.geo[B.geo];
  const int64_t id = cell ? ((int64_t)k * B.n[1] + j) * B.n[0] + i : 0;
  const int64_t t = cell ? cen_index(G, k, j, i) : 0;
  const bool active =
      cell && !(P.metric == PANGU_METRIC_CKS && ldc32(G.cen, (int64_t)CG_RAD * G.ncen, (uint32_t)t) < P.r_excise);
  int prev = 0;
  if (active && S.iters) prev = glb(S.iters[b])[id];
  const bool gated = active && S.iters && prev >= S.gate;
  double x0 = -1.0;
  __shared__ MlpLds lds;
  // every thread of the workgroup reaches the vote (no early return above)
  if (__syncthreads_or(gated)) {
    mlp_stage(W, lds, threadIdx.x, blockDim.x);
    KastaunIn kin;
    const int setup = active ? learned_setup<HB, HK>(C, F, G, t, b, id, kin) : -1;
    __syncthreads();
    const bool want = setup == 0;
    if (__ballot(gated && want)) {  // uniform over the wavefront: all 64 lanes run the network
      float x[5] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
      if (want) {
        float f[5];
        c2p_features(kin, f);
#pragma unroll
        for (int q = 0; q < 5; ++q) x[q] = bf2f(f2bf((f[q] - W.mean[q]) / W.dev[q]));
      }
      const float z = wave_mlp(lds, W.b4[0], x, lane);
      if (want) x0 = (double)(1.0f / (1.0f + expf(-z)));
    }
  }
  if (!cell) return;
  recover_cell<2, HB, HK, HEAT>(C, F, S, b, k, j, i, x0);
}

// ---------------------------------------------------------- FixRecovery
// recovery_fixer.cc:216-285 with the serial (lexicographic, in-place)
// semantics of a single-threaded host build.  Phase A fixes failed cells
// with no failed 8-neighbour in parallel (nobody else reads or writes their
// stencils); cells touching another failed cell are left for phase B,
// which walks each (block, k) plane in lexicographic order.
__device__ __forceinline__ void fix_cell(double *const *pp, int np, int64_t id, int64_t jp,
                                         const double *fl) {
  const int pf1 = (int)fl[id + jp - 1], pf2 = (int)fl[id + jp], pf3 = (int)fl[id + jp + 1];
  const int pf4 = (int)fl[id + 1];
  const int pf5 = (int)fl[id - jp + 1], pf6 = (int)fl[id - jp], pf7 = (int)fl[id - jp - 1];
  const int pf8 = (int)fl[id - 1];
  for (int q = 0; q < np; ++q) {
    double *a = pp[q];
    if (pf2 && pf4 && pf6 && pf8)
      a[id] = 0.25 * (a[id + jp] + a[id - jp] + a[id - 1] + a[id + 1]);
    else if (pf1 && pf3 && pf5 && pf7)
      a[id] = 0.25 * (a[id + jp + 1] + a[id + jp - 1] + a[id - jp + 1] + a[id - jp - 1]);
    else
      a[id] = 0.125 * (a[id + jp - 1] + a[id + jp] + a[id + jp + 1] + a[id + 1] + a[id - jp + 1] +
                       a[id - jp] + a[id - jp - 1] + a[id - 1]);
  }
  pp[2][id] = 0.;
  pp[3][id] = 0.;
  pp[4][id] = 0.;
}

__global__ __launch_bounds__(256) void k_fixrec_a(KCtx C, KFields F, double *const *flag,
                                                  int *cluster, Range3 R) {
  int b, k, j, i;
  if (!decode_tile(C, R, b, k, j, i)) return;
  const DevBlock &B = C.blk[b];
  const int kk = k - B.s[2];
  const int64_t id = ((int64_t)k * B.n[1] + j) * B.n[0] + i;
  double *fl = flag[b];
  if (fl[id] != 0) return;
  if (C.ph.metric == PANGU_METRIC_CKS) {
    const DevGeo &G = C.geo[B.geo];
    if (G.cen[CG_RAD * G.ncen + cen_index(G, k, j, i)] < C.ph.r_excise) return;
  }
  const int64_t jp = (B.s[1] != B.e[1]) ? (int64_t)B.n[0] : 0;
  const bool isolated = fl[id + jp - 1] != 0 && fl[id + jp] != 0 && fl[id + jp + 1] != 0 &&
                        fl[id + 1] != 0 && fl[id - jp + 1] != 0 && fl[id - jp] != 0 &&
                        fl[id - jp - 1] != 0 && fl[id - 1] != 0;
  if (!isolated) {
    atomicOr(&cluster[C.pfirst[b] + kk], 1);
    return;
  }
  fix_cell(F.prim + b * C.np, C.np, id, jp, fl);
  fl[id] = 1;
}

// Phase B: one wavefront per interior (block, k) plane that holds clustered
// cells (workgroup index = plane index).
__global__ __launch_bounds__(64) void k_fixrec_b(KCtx C, KFields F, double *const *flag,
                                                 int *cluster) {
  const int b = C.pblock[blockIdx.x];
  const int kk = (int)blockIdx.x - C.pfirst[b];
  if (cluster[blockIdx.x] == 0) return;
  const DevBlock &B = C.blk[b];
  const int k = B.s[2] + kk;
  const int64_t jp = (B.s[1] != B.e[1]) ? (int64_t)B.n[0] : 0;
  double *fl = flag[b];
  const int lane = threadIdx.x;
  for (int j = B.s[1]; j <= B.e[1]; ++j) {
    for (int i0 = B.s[0]; i0 <= B.e[0]; i0 += 64) {
      const int i = i0 + lane;
      const int64_t id = ((int64_t)k * B.n[1] + j) * B.n[0] + i;
      bool bad = (i <= B.e[0]) && fl[id] == 0;
      if (bad && C.ph.metric == PANGU_METRIC_CKS) {
        const DevGeo &G = C.geo[B.geo];
        if (G.cen[CG_RAD * G.ncen + cen_index(G, k, j, i)] < C.ph.r_excise) bad = false;
      }
      uint64_t m = __ballot(bad);
      while (m) {
        const int l = __ffsll((unsigned long long)m) - 1;
        m &= m - 1;
        if (lane == 0) {
          const int64_t idl = id - lane + l;
          fix_cell(F.prim + b * C.np, C.np, idl, jp, fl);
          fl[idl] = 1;
        }
        __syncthreads();
      }
    }
  }
  if (lane == 0) cluster[blockIdx.x] = 0;
}

// ------------------------------------------------- ghost exchange and BCs
// Ghost cell g of a block, numbered over three slabs: the k-ghost planes
// (full j, i extent), then the j-ghost rows of interior planes (full i),
// then the i-ghost cells of interior rows.  Every ghost cell appears once,
// with i fastest inside each slab.
__device__ __forceinline__ void ghost_cell(const DevBlock &B, int g, int &k, int &j, int &i) {
  const int n0 = B.n[0], n1 = B.n[1], n2 = B.n[2];
  const int a2 = B.s[2] + (n2 - 1 - B.e[2]);
  const int A = a2 * n1 * n0;
  if (g < A) {
    const int kk = g / (n1 * n0), r = g - kk * n1 * n0;
    k = kk < B.s[2] ? kk : B.e[2] + 1 + (kk - B.s[2]);
    j = r / n0;
    i = r - j * n0;
    return;
  }
  g -= A;
  const int b1 = B.s[1] + (n1 - 1 - B.e[1]);
  const int nk = B.e[2] - B.s[2] + 1;
  const int Bc = nk * b1 * n0;
  if (g < Bc) {
    const int kk = g / (b1 * n0), r = g - kk * b1 * n0;
    const int jj = r / n0;
    k = B.s[2] + kk;
    j = jj < B.s[1] ? jj : B.e[1] + 1 + (jj - B.s[1]);
    i = r - jj * n0;
    return;
  }
  g -= Bc;
  const int nj = B.e[1] - B.s[1] + 1;
  const int c0 = B.s[0] + (n0 - 1 - B.e[0]);
  const int kk = g / (nj * c0), r = g - kk * nj * c0;
  const int jj = r / c0, ii = r - jj * c0;
  k = B.s[2] + kk;
  j = B.s[1] + jj;
  i = ii < B.s[0] ? ii : B.e[0] + 1 + (ii - B.s[0]);
}

// Same-level exchange: a ghost cell of block b takes the interior value of
// the neighbour in its offset direction (periodic wrap encoded in nbr).
// One thread per ghost cell.
__global__ __launch_bounds__(256) void k_exchange(KCtx C, KFields F, double *const *flag,
                                                  const int *nbr, Items I) {
  int b, g;
  decode_item(I, b, g);
  const DevBlock &B = C.blk[b];
  const int interior = (B.e[0] - B.s[0] + 1) * (B.e[1] - B.s[1] + 1) * (B.e[2] - B.s[2] + 1);
  if (g >= (int)B.ncell - interior) return;
  int k, j, i;
  ghost_cell(B, g, k, j, i);
  const int o0 = (i < B.s[0]) ? -1 : (i > B.e[0] ? 1 : 0);
  const int o1 = (j < B.s[1]) ? -1 : (j > B.e[1] ? 1 : 0);
  const int o2 = (k < B.s[2]) ? -1 : (k > B.e[2] ? 1 : 0);
  if ((o0 | o1 | o2) == 0) return;
  const int nb = nbr[b * 27 + (o2 + 1) * 9 + (o1 + 1) * 3 + (o0 + 1)];
  if (nb < 0) return;
  const DevBlock &N = C.blk[nb];
  const int si = i - o0 * (B.e[0] - B.s[0] + 1);
  const int sj = j - o1 * (B.e[1] - B.s[1] + 1);
  const int sk = k - o2 * (B.e[2] - B.s[2] + 1);
  const int64_t dst = ((int64_t)k * B.n[1] + j) * B.n[0] + i;
  const int64_t src = ((int64_t)sk * N.n[1] + sj) * N.n[0] + si;
  for (int q = 0; q < C.np; ++q) glb(F.prim[b * C.np + q])[dst] = glb(F.prim[nb * C.np + q])[src];
  if (flag) glb(flag[b])[dst] = glb(flag[nb])[src];
}

// Cell g of the two ghost slabs normal to d (side 0 below s[d], side 1 above
// e[d]), each ng layers deep and spanning the full ghosted extent of the
// other dims; i is fastest except for d = 0, where the layer index is.
__device__ __forceinline__ void bc_cell(const DevBlock &B, int d, int g, int &side, int &k, int &j,
                                        int &i) {
  const int ng = B.s[d];
  int ix[3];
  if (d == 0) {
    const int na = B.n[1], nb = B.n[2], per = ng * na * nb;
    side = g / per;
    const int r = g - side * per, kk = r / (na * ng), r2 = r - kk * na * ng, jj = r2 / ng;
    ix[0] = r2 - jj * ng;
    ix[1] = jj;
    ix[2] = kk;
  } else if (d == 1) {
    const int na = B.n[0], nb = B.n[2], per = ng * na * nb;
    side = g / per;
    const int r = g - side * per, kk = r / (ng * na), r2 = r - kk * ng * na, l = r2 / na;
    ix[0] = r2 - l * na;
    ix[1] = l;
    ix[2] = kk;
  } else {
    const int na = B.n[0], nb = B.n[1], per = ng * na * nb;
    side = g / per;
    const int r = g - side * per, l = r / (nb * na), r2 = r - l * nb * na, jj = r2 / na;
    ix[0] = r2 - jj * na;
    ix[1] = jj;
    ix[2] = l;
  }
  if (side == 1) ix[d] = B.e[d] + 1 + ix[d];
  i = ix[0];
  j = ix[1];
  k = ix[2];
}

// Physical boundaries normal to dimension d: outflow copies the edge cell;
// reflecting mirrors and flips the d-component of u~ and B.  Covers the full
// extent of the other dims, so X1, X2, X3 launches in turn fill corners.
// One thread per cell of the two ghost slabs of each block with a physical
// face normal to d; B.bcf says which faces are physical and their kind.
__global__ __launch_bounds__(256) void k_bc(KCtx C, KFields F, double *const *flag, int d, Items I) {
  int b, g;
  decode_item(I, b, g);
  const DevBlock &B = C.blk[b];
  const int per_block = 2 * B.s[d] *
                        (d == 0 ? B.n[1] * B.n[2] : (d == 1 ? B.n[0] * B.n[2] : B.n[0] * B.n[1]));
  if (g >= per_block) return;
  int side, k, j, i;
  bc_cell(B, d, g, side, k, j, i);
  const int kind = B.bcf[d][side];
  if (kind < 0) return;
  const int ix[3] = {i, j, k};
  int sx[3] = {i, j, k};
  if (kind == PANGU_BC_OUTFLOW)
    sx[d] = side == 0 ? B.s[d] : B.e[d];
  else
    sx[d] = side == 0 ? (2 * B.s[d] - 1 - ix[d]) : (2 * B.e[d] + 1 - ix[d]);
  const int64_t dst = ((int64_t)k * B.n[1] + j) * B.n[0] + i;
  const int64_t src = ((int64_t)sx[2] * B.n[1] + sx[1]) * B.n[0] + sx[0];
  const int iBX = C.iBX;
  for (int q = 0; q < C.np; ++q) {
    double v = glb(F.prim[b * C.np + q])[src];
    if (kind == PANGU_BC_REFLECT && (q == 2 + d || (iBX >= 0 && q == iBX + d))) v = -v;
    glb(F.prim[b * C.np + q])[dst] = v;
  }
  if (flag) glb(flag[b])[dst] = glb(flag[b])[src];
}

// --------------------------------------------------------------- timestep
// Per block and active direction the maximum face speed over interior
// cells; min_cells(cfl dx / c) equals cfl dx / max_cells(c) because IEEE
// division is monotone.  NaN speeds are skipped, as the reference's
// `if (dt_min > dt)` reduction skips NaN.  One workgroup per item (block,
// interior plane, chunk of 256 * PER cells of the plane) writes its maxima
// to red[item*3 + d]; no atomics, so the result is deterministic.
template <int PER>
__global__ __launch_bounds__(256) void k_dt_reduce(KCtx C, double *const *alfven, double *red) {
  __shared__ double sm[4][3];
  // thread x takes cells x, x + 256, ... of the chunk, all loads in flight
  // together (the maximum does not depend on the order)
  constexpr int kPer = PER, kDtChunk = 256 * PER;
  const int b = C.dtblock[blockIdx.x];
  const int local = (int)blockIdx.x - C.dtfirst[b];
  const DevBlock &B = C.blk[b];
  const int ni = B.e[0] - B.s[0] + 1, nj = B.e[1] - B.s[1] + 1;
  const int nch = (ni * nj + kDtChunk - 1) / kDtChunk;
  const int kk = local / nch, ch = local - kk * nch;
  const int k = B.s[2] + kk;
  const int t0 = ch * kDtChunk, t1 = (t0 + kDtChunk < ni * nj) ? t0 + kDtChunk : ni * nj;
  // (j, i) of cell t, stepped by 256 cells without a division per step
  const int dj = 256 / ni, di = 256 - dj * ni;
  int t = t0 + (int)threadIdx.x;
  int j = t / ni, i = t - j * ni;
  // steps holding a cell for some thread (the same for the whole workgroup:
  // a small plane skips the rest)
  const int nr = (t1 - t0 + 255) / 256;
  double v[kPer][3];
#pragma unroll
  for (int r = 0; r < kPer; ++r) {
    const int64_t id = ((int64_t)k * B.n[1] + (B.s[1] + j)) * B.n[0] + (B.s[0] + i);
#pragma unroll
    for (int d = 0; d < 3; ++d)
      v[r][d] = (r < nr && t < t1 && B.n[d] != 1) ? glb(alfven[b * 3 + d])[id] : 0.0;
    t += 256;
    i += di;
    j += dj;
    if (i >= ni) {
      i -= ni;
      ++j;
    }
  }
  double c[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int r = 0; r < kPer; ++r)
#pragma unroll
    for (int d = 0; d < 3; ++d)
      if (v[r][d] > c[d]) c[d] = v[r][d];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int d = 0; d < 3; ++d) {
    double v = c[d];
    for (int off = 32; off > 0; off >>= 1) {
      const double o = __shfl_xor(v, off, 64);
      v = (o > v) ? o : v;
    }
    if (lane == 0) sm[w][d] = v;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    const int d = threadIdx.x;
    double v = sm[0][d];
    for (int q = 1; q < 4; ++q) v = (sm[q][d] > v) ? sm[q][d] : v;
    red[(int64_t)blockIdx.x * 3 + d] = v;
  }
}

// Harmonic combination per block (timestep_estimation.cc:94-98), minimum
// over the blocks of one mesh, then dt control dt_next = min(dt_factor * dt,
// estimate) and time += dt.  One workgroup per mesh m: state + 2m = {dt,
// time}; est_out (optional) + m gets the estimate.
__global__ __launch_bounds__(256) void k_dt_final(KCtx C, const double *red, double *state,
                                                  double dt_factor, double *est_out) {
  __shared__ double sm[256];
  const int m = blockIdx.x;
  double mn = DBL_MAX;
  for (int b = C.mfirst[m] + (int)threadIdx.x; b < C.mfirst[m + 1]; b += blockDim.x) {
    const DevBlock &B = C.blk[b];
    double q3[3] = {DBL_MAX, DBL_MAX, DBL_MAX};
    for (int d = 0; d < 3; ++d) {
      if (B.n[d] == 1) continue;
      // the block's items, 16 independent loads at a time
      double cmax = 0.0;
      const int p1 = C.dtfirst[b + 1];
      for (int p0 = C.dtfirst[b]; p0 < p1; p0 += 16) {
        double w[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) w[u] = (p0 + u < p1) ? red[(int64_t)(p0 + u) * 3 + d] : 0.0;
#pragma unroll
        for (int u = 0; u < 16; ++u) cmax = (w[u] > cmax) ? w[u] : cmax;
      }
      const double q = C.ph.cfl * B.dx[d] / cmax;
      q3[d] = (q < DBL_MAX) ? q : DBL_MAX;
    }
    const double t = 1.0 / (1.0 / q3[0] + 1.0 / q3[1] + 1.0 / q3[2]);
    if (t < mn) mn = t;
  }
  sm[threadIdx.x] = mn;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      const double o = sm[threadIdx.x + s];
      if (o < sm[threadIdx.x]) sm[threadIdx.x] = o;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double est = sm[0];
    if (est_out) est_out[m] = est;
    if (state) {
      double *st = state + 2 * m;
      const double dt = st[0];
      st[1] = st[1] + dt;
      const double grown = dt_factor * dt;
      st[0] = (est < grown) ? est : grown;
    }
  }
}

// ===================================================================== host
// Grid of a tile launch, padded to a multiple of the 8 XCDs.
static dim3 tile_grid(const Range3 &R) { return dim3((unsigned)(((R.ntiles + 7) / 8) * 8)); }
static const dim3 kTile(TILE, 1, 1);

// Grid of a persistent kernel: the workgroups the GPU holds at once (at
// most ntiles), so every workgroup runs from the start.
template <typename K>
static int resident_groups(K kernel, int threads, int ntiles) {
  int dev = 0, per_cu = 0, ncu = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0) != hipSuccess ||
      hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return 1;
  const int64_t n = (int64_t)(per_cu > 0 ? per_cu : 1) * (ncu > 0 ? ncu : 1);
  return (int)(n < ntiles ? n : ntiles);
}

// Range of a launch of kernel family fam over base range id (0 interior,
// 1 entire, 2..4 the faces, 5 CT) in that family's tile size and order.
static Range3 family_range(const LaunchCtx &L, int id, const Range3 &base, int fam) {
  const int t = L.tile[fam];
  Range3 R = (t == 128) ? L.small[id][0] : (t == 64) ? L.small[id][1] : base;
  R.order = L.order[fam];
  return R;
}

#define PG_CHECK_LAUNCH() (hipGetLastError() == hipSuccess ? 0 : -2)
#define TICK(name) \
  if (L.clock) (void)L.clock->mark(name, s)

hipError_t KernelClock::mark(const char *name, hipStream_t s) {
  if (n == events.size()) {
    hipEvent_t e;
    const hipError_t r = hipEventCreate(&e);
    if (r != hipSuccess) return r;
    events.push_back(e);
  }
  if (n > 0) {
    if (names.size() < n) names.resize(n);
    names[n - 1] = name;
  }
  return hipEventRecord(events[n++], s);
}

KernelClock::~KernelClock() {
  for (hipEvent_t e : events) (void)hipEventDestroy(e);
}

int launch_fix_cons(const LaunchCtx &L, const KFields &F, uint8_t *const *mask, bool fix, bool cons,
                    hipStream_t s) {
  const Range3 R = fix ? family_range(L, 1, L.entire, PANGU_KF_FIX_CONS)
                       : family_range(L, 0, L.interior, PANGU_KF_FIX_CONS);
  if (R.ntiles == 0) return 0;
  const dim3 g = tile_grid(R), wg(R.tile);
  const bool hb = L.k.iBX >= 0, hk = L.k.iKEL >= 0;
#define FC(FX, CS, HB_, HK_) hipLaunchKernelGGL((k_fix_cons<FX, CS, HB_, HK_>), g, wg, 0, s, L.k, F, mask, R)
#define FC2(FX, CS)                      \
  if (hb && hk) FC(FX, CS, true, true);  \
  else if (hb) FC(FX, CS, true, false);  \
  else if (hk) FC(FX, CS, false, true);  \
  else FC(FX, CS, false, false);
  if (fix && cons) {
    FC2(true, true)
  } else if (fix) {
    FC2(true, false)
  } else {
    FC2(false, true)
  }
#undef FC2
#undef FC
  TICK(fix && cons ? "fix_cons" : fix ? "fix_primitive" : "conservative");
  return PG_CHECK_LAUNCH();
}

template <int DIR>
static void flux_dir(const LaunchCtx &L, const KFields &F, double *const *alfven, hipStream_t s) {
  const bool shared = DIR == 1 && L.share_x1;
  Range3 R = shared ? L.x1cells : L.face[DIR - 1];
  R.order = L.order[PANGU_KF_FLUX_X1 + DIR - 1];
  constexpr int MD = DIR >= 2 ? DIR - 2 : 0;
  const bool march = DIR >= 2 && L.march_depth[MD] > 0;
  if (march) {
    R = L.march[MD];
    R.order = L.order[PANGU_KF_FLUX_X1 + DIR - 1];
  }
  const bool pair = !shared && !march && ((L.pair_flux >> (DIR - 1)) & 1);
  // the one-lane sweep in the family's tile size (the other forms decode
  // 256-cell tiles)
  if (!shared && !march && !pair) R = family_range(L, 1 + DIR, L.face[DIR - 1], PANGU_KF_FLUX_X1 + DIR - 1);
  if (R.ntiles == 0) return;
  const dim3 g = tile_grid(R), wg(shared || march || pair ? TILE : R.tile);
  const dim3 g2((unsigned)(((2 * R.ntiles + 7) / 8) * 8));
  const bool hb = L.k.iBX >= 0, hk = L.k.iKEL >= 0;
  const int lim = L.limiter, rs = L.riemann;
#define FL(LIM, RS, HB_, HK_)                                                                      \
  if (shared)                                                                                     \
    hipLaunchKernelGGL((k_flux_x1s<LIM, RS, HB_, HK_>), g, kTile, 0, s, L.k, F, alfven, R);       \
  else if (march)                                                                                 \
    hipLaunchKernelGGL((k_flux_march<(DIR > 1 ? DIR : 2), LIM, RS, HB_, HK_>), g, kTile, 0, s, L.k, F, \
                       alfven, R);                                                                \
  else if (pair)                                                                                  \
    hipLaunchKernelGGL((k_flux_pair<DIR, LIM, RS, HB_, HK_>), g2, kTile, 0, s, L.k, F, alfven, R); \
  else                                                                                            \
    hipLaunchKernelGGL((k_flux<DIR, LIM, RS, HB_, HK_>), g, wg, 0, s, L.k, F, alfven, R)
#define FL2(LIM, RS)                    \
  if (hb && hk) FL(LIM, RS, true, true); \
  else if (hb) FL(LIM, RS, true, false); \
  else if (hk) FL(LIM, RS, false, true); \
  else FL(LIM, RS, false, false);
  if (lim == PANGU_LIMITER_PPM4 && rs == PANGU_RIEMANN_LAXF) {
    FL2(PANGU_LIMITER_PPM4, PANGU_RIEMANN_LAXF)
  } else if (lim == PANGU_LIMITER_PPM4) {
    FL2(PANGU_LIMITER_PPM4, PANGU_RIEMANN_HLL)
  } else if (rs == PANGU_RIEMANN_LAXF) {
    FL2(PANGU_LIMITER_MC, PANGU_RIEMANN_LAXF)
  } else {
    FL2(PANGU_LIMITER_MC, PANGU_RIEMANN_HLL)
  }
#undef FL2
#undef FL
  TICK(DIR == 1 ? "flux_x1" : DIR == 2 ? "flux_x2" : "flux_x3");
}

int launch_fluxes_23(const LaunchCtx &L, const KFields &F, double *const *alfven, hipStream_t s) {
  flux_dir<2>(L, F, alfven, s);
  flux_dir<3>(L, F, alfven, s);
  return PG_CHECK_LAUNCH();
}

template <int LIM, int RS, bool HB, bool HK>
static void launch_fx(const LaunchCtx &L, const Range3 &R, const KFields &F, uint8_t *const *mask,
                      double *const *alfven, hipStream_t s) {
  hipLaunchKernelGGL((k_fix_flux_x1<LIM, RS, HB, HK>), tile_grid(R), kTile, 0, s, L.k, F, mask, alfven, R);
}
template <int LIM, int RS>
static void launch_fx_fields(const LaunchCtx &L, const Range3 &R, const KFields &F, uint8_t *const *mask,
                             double *const *alfven, hipStream_t s) {
  const bool hb = L.k.iBX >= 0, hk = L.k.iKEL >= 0;
  if (hb && hk)
    launch_fx<LIM, RS, true, true>(L, R, F, mask, alfven, s);
  else if (hb)
    launch_fx<LIM, RS, true, false>(L, R, F, mask, alfven, s);
  else if (hk)
    launch_fx<LIM, RS, false, true>(L, R, F, mask, alfven, s);
  else
    launch_fx<LIM, RS, false, false>(L, R, F, mask, alfven, s);
}

int launch_fix_flux_x1(const LaunchCtx &L, const KFields &F, uint8_t *const *mask, double *const *alfven,
                       hipStream_t s) {
  Range3 R = L.fixx1;
  R.order = L.order[PANGU_KF_FIX_CONS];
  if (R.ntiles == 0 || R.tile + 8 != TILE) return -1;
  if (L.limiter == PANGU_LIMITER_PPM4) {
    if (L.riemann == PANGU_RIEMANN_HLL)
      launch_fx_fields<PANGU_LIMITER_PPM4, PANGU_RIEMANN_HLL>(L, R, F, mask, alfven, s);
    else
      launch_fx_fields<PANGU_LIMITER_PPM4, PANGU_RIEMANN_LAXF>(L, R, F, mask, alfven, s);
  } else {
    if (L.riemann == PANGU_RIEMANN_HLL)
      launch_fx_fields<PANGU_LIMITER_MC, PANGU_RIEMANN_HLL>(L, R, F, mask, alfven, s);
    else
      launch_fx_fields<PANGU_LIMITER_MC, PANGU_RIEMANN_LAXF>(L, R, F, mask, alfven, s);
  }
  TICK("fix_cons_flux_x1");
  return PG_CHECK_LAUNCH();
}

int launch_fluxes(const LaunchCtx &L, const KFields &F, double *const *alfven, hipStream_t s) {
  flux_dir<1>(L, F, alfven, s);
  flux_dir<2>(L, F, alfven, s);
  flux_dir<3>(L, F, alfven, s);
  return PG_CHECK_LAUNCH();
}

int launch_ct(const LaunchCtx &L, const KFields &F, double *const *emf, hipStream_t s) {
  const Range3 R = family_range(L, 5, L.ct, PANGU_KF_CT);
  if (R.ntiles == 0) return 0;
  const dim3 g = tile_grid(R), wg(R.tile);
  hipLaunchKernelGGL(k_emf, g, wg, 0, s, L.k, F, emf, R);
  TICK("ct_emf");
  hipLaunchKernelGGL(k_ct_faces, g, wg, 0, s, L.k, F, emf, R);
  TICK("ct_faces");
  return PG_CHECK_LAUNCH();
}

int launch_div_update(const LaunchCtx &L, int mode, const KFields &F0, const KFields &Finit,
                      const KFields &F1, double *const *dudt, bool ct_fused,
                      const double *dt_dev, double beta, double bdt_host, hipStream_t s) {
  const Range3 R = family_range(L, 0, L.interior, PANGU_KF_DIV_UPDATE);
  if (R.ntiles == 0) return 0;
  const dim3 g = tile_grid(R), wg(R.tile);
  const bool hb = L.k.iBX >= 0, hk = L.k.iKEL >= 0;
  const bool src = L.metric != PANGU_METRIC_MINKOWSKI;
  // CT fused into the divergence (modes 4 and 5 with a field)
  const bool ctf = ct_fused && hb && (mode == 4 || mode == 5);
#define DU(M, SR, HB_, HK_, CT_)                                                               \
  hipLaunchKernelGGL((k_div_update<M, SR, HB_, HK_, CT_>), g, wg, 0, s, L.k, F0, Finit, F1, \
                     dudt, dt_dev, beta, bdt_host, R)
#define DU2(M, SR)                                      \
  if (hb && hk) {                                       \
    if ((M) == 4 && ctf) DU(M, SR, true, true, (M) == 4); \
    else DU(M, SR, true, true, false);                  \
  } else if (hb) {                                      \
    if ((M) == 4 && ctf) DU(M, SR, true, false, (M) == 4); \
    else DU(M, SR, true, false, false);                 \
  } else if (hk) DU(M, SR, false, true, false);         \
  else DU(M, SR, false, false, false);
  switch (mode) {
    case 1: DU2(1, false) break;
    case 2: DU2(2, false) break;
    case 3:
      if (src) {
        DU2(3, true)
      }
      break;
    case 5: DU2(4, false) break;
    default:
      if (src) {
        DU2(4, true)
      } else {
        DU2(4, false)
      }
  }
#undef DU2
#undef DU
  if (mode != 3 || src)
    TICK(mode == 1   ? "flux_divergence"
         : mode == 2 ? "update"
         : mode == 3 ? "source"
         : mode == 5 ? "div_update"
                     : "div_update_source");
  return PG_CHECK_LAUNCH();
}

template <bool HB, bool HK, bool HEAT>
static void launch_learned(const LaunchCtx &L, const KFields &F, const SharedTabs &S, const MlpFrag &W,
                           const Range3 &R, dim3 g, dim3 wg, hipStream_t s, const UpdArgs *upd) {
  const UpdArgs A = upd ? *upd : UpdArgs{};
  const int u = !upd ? 0 : (L.metric != PANGU_METRIC_MINKOWSKI ? 2 : 1);
  if (W.kind == 0) {
    if (u == 2)
      hipLaunchKernelGGL((k_recover_learned<HB, HK, HEAT, 2>), g, wg, 0, s, L.k, F, S, W, R, A);
    else if (u == 1)
      hipLaunchKernelGGL((k_recover_learned<HB, HK, HEAT, 1>), g, wg, 0, s, L.k, F, S, W, R, A);
    else
      hipLaunchKernelGGL((k_recover_learned<HB, HK, HEAT, 0>), g, wg, 0, s, L.k, F, S, W, R, A);
  } else {
    if (u == 2)
      hipLaunchKernelGGL((k_recover_learned_vsq<HB, HK, HEAT, 2>), g, wg, 0, s, L.k, F, S, W, R, A);
    else if (u == 1)
      hipLaunchKernelGGL((k_recover_learned_vsq<HB, HK, HEAT, 1>), g, wg, 0, s, L.k, F, S, W, R, A);
    else
      hipLaunchKernelGGL((k_recover_learned_vsq<HB, HK, HEAT, 0>), g, wg, 0, s, L.k, F, S, W, R, A);
  }
}

int launch_recover(const LaunchCtx &L, const KFields &F, const SharedTabs &S, int rec, bool heat,
                   const MlpFrag *W, hipStream_t s, const UpdArgs *upd) {
  // a fused update needs the learned kernel of the scheme (k_update_recover
  // serves the plain ones) and the range of the update
  if (upd && !(W && ((rec == 1 && W->kind == 0) || (rec == 2 && W->kind == 1)))) return -1;
  const Range3 R = upd ? family_range(L, 0, L.interior, PANGU_KF_DIV_UPDATE)
                       : family_range(L, 0, L.interior, PANGU_KF_RECOVERY);
  if (R.ntiles == 0) return 0;
  const dim3 g = tile_grid(R), wg(R.tile);
  const bool hb = L.k.iBX >= 0, hk = L.k.iKEL >= 0;
  const bool persist = L.persist_recovery && L.queue;
#define RC(R_, HB_, HK_, HE_)                                                                        \
  if ((R_ == 1 && W && W->kind == 0) || (R_ == 2 && W && W->kind == 1))                              \
    launch_learned<HB_, HK_, HE_>(L, F, S, *W, R, g, wg, s, upd);                                     \
  else if (persist)                                                                                   \
    hipLaunchKernelGGL((k_recover_queue<R_, HB_, HK_, HE_>),                                          \
                       dim3(resident_groups(k_recover_queue<R_, HB_, HK_, HE_>, R.tile, R.ntiles)), wg, \
                       0, s, L.k, F, S, R, L.queue);                                                  \
  else                                                                                                \
    hipLaunchKernelGGL((k_recover<R_, HB_, HK_, HE_>), g, wg, 0, s, L.k, F, S, R)
#define RC2(R_, HE_)                     \
  if (hb && hk) RC(R_, true, true, HE_);  \
  else if (hb) RC(R_, true, false, HE_);  \
  else if (hk) RC(R_, false, true, HE_);  \
  else RC(R_, false, false, HE_);
  const bool h = heat && hk;
  if (rec == 0) {
    if (h) { RC2(0, true) } else { RC2(0, false) }
  } else if (rec == 1) {
    if (h) { RC2(1, true) } else { RC2(1, false) }
  } else if (rec == 2) {
    if (h) { RC2(2, true) } else { RC2(2, false) }
  } else if (h) {
    RC2(-1, true)
  }
#undef RC2
#undef RC
  if (rec >= 0 || h) TICK(rec < 0 ? "heating" : h ? "recovery_heating" : "recovery");
  return PG_CHECK_LAUNCH();
}

// Arguments of one k_update_recover launch, and the dispatch over its
// compile-time options: field (with CT formed in the update), K_e, heating
// and the recovery scheme.
struct URArgs {
  const KCtx &k;
  const KFields &F0, &Finit, &F1;
  const SharedTabs &S;
  const double *dt_dev;
  double beta, bdt_host;
  const Range3 &R;
};
template <bool SRC, bool HB, bool HK, int REC, bool HEAT>
static void launch_ur(const URArgs &a, dim3 g, dim3 wg, hipStream_t s) {
  hipLaunchKernelGGL((k_update_recover<SRC, HB, HK, HB, REC, HEAT>), g, wg, 0, s, a.k, a.F0, a.Finit,
                     a.F1, a.S, a.dt_dev, a.beta, a.bdt_host, a.R);
}
template <bool SRC, int REC>
static void launch_ur_fields(const URArgs &a, dim3 g, dim3 wg, hipStream_t s, bool hb, bool hk, bool h) {
  if (!hb)
    launch_ur<SRC, false, false, REC, false>(a, g, wg, s);
  else if (h)
    launch_ur<SRC, true, true, REC, true>(a, g, wg, s);
  else if (hk)
    launch_ur<SRC, true, true, REC, false>(a, g, wg, s);
  else
    launch_ur<SRC, true, false, REC, false>(a, g, wg, s);
}
template <bool SRC>
static void launch_ur_rec(const URArgs &a, dim3 g, dim3 wg, hipStream_t s, int rec, bool hb, bool hk,
                          bool h) {
  if (rec == 0)
    launch_ur_fields<SRC, 0>(a, g, wg, s, hb, hk, h);
  else if (rec == 1)
    launch_ur_fields<SRC, 1>(a, g, wg, s, hb, hk, h);
  else
    launch_ur_fields<SRC, 2>(a, g, wg, s, hb, hk, h);
}

int launch_update_recover(const LaunchCtx &L, const KFields &F0, const KFields &Finit,
                          const KFields &F1, const SharedTabs &S, int rec, bool heat, bool ct_fused,
                          const double *dt_dev, double beta, double bdt_host, hipStream_t s) {
  const Range3 R = family_range(L, 0, L.interior, PANGU_KF_DIV_UPDATE);
  if (R.ntiles == 0) return 0;
  const dim3 g = tile_grid(R), wg(R.tile);
  const bool hb = L.k.iBX >= 0, hk = L.k.iKEL >= 0, h = heat && hk;
  const bool src = L.metric != PANGU_METRIC_MINKOWSKI;
  if ((hb && !ct_fused) || (!hb && hk) || rec < 0 || rec > 2) return -1;
  const URArgs a{L.k, F0, Finit, F1, S, dt_dev, beta, bdt_host, R};
  if (src)
    launch_ur_rec<true>(a, g, wg, s, rec, hb, hk, h);
  else
    launch_ur_rec<false>(a, g, wg, s, rec, hb, hk, h);
  TICK("update_recovery");
  return PG_CHECK_LAUNCH();
}

int launch_fix_recovery(const LaunchCtx &L, const KFields &F, double *const *flag, int *cluster,
                        hipStream_t s) {
  const Range3 &R = L.interior;
  if (R.ntiles == 0) return 0;
  hipLaunchKernelGGL(k_fixrec_a, tile_grid(R), kTile, 0, s, L.k, F, flag, cluster, R);
  TICK("fix_recovery_a");
  hipLaunchKernelGGL(k_fixrec_b, dim3(L.k.nplanes), dim3(64), 0, s, L.k, F, flag, cluster);
  TICK("fix_recovery_b");
  return PG_CHECK_LAUNCH();
}

int launch_exchange(const LaunchCtx &L, const KFields &F, double *const *flag, const int *nbr,
                    hipStream_t s) {
  if (L.ghosts.ntiles > 0)
    hipLaunchKernelGGL(k_exchange, dim3(L.ghosts.ntiles), kTile, 0, s, L.k, F, flag, nbr, L.ghosts);
  TICK("exchange");
  return PG_CHECK_LAUNCH();
}

int launch_bc(const LaunchCtx &L, const KFields &F, double *const *flag, hipStream_t s) {
  for (int d = 0; d < 3; ++d) {
    if (L.bcs[d].ntiles == 0) continue;
    hipLaunchKernelGGL(k_bc, dim3(L.bcs[d].ntiles), kTile, 0, s, L.k, F, flag, d, L.bcs[d]);
    TICK(d == 0 ? "bc_x1" : d == 1 ? "bc_x2" : "bc_x3");
  }
  return PG_CHECK_LAUNCH();
}

int launch_dt(const LaunchCtx &L, double *const *alfven, double *red, double *state,
              double dt_factor, double *est_out, hipStream_t s) {
  if (L.k.dtper == 1)
    hipLaunchKernelGGL(k_dt_reduce<1>, dim3(L.k.ndt), dim3(256), 0, s, L.k, alfven, red);
  else if (L.k.dtper == 4)
    hipLaunchKernelGGL(k_dt_reduce<4>, dim3(L.k.ndt), dim3(256), 0, s, L.k, alfven, red);
  else
    hipLaunchKernelGGL(k_dt_reduce<16>, dim3(L.k.ndt), dim3(256), 0, s, L.k, alfven, red);
  TICK("dt_reduce");
  hipLaunchKernelGGL(k_dt_final, dim3(L.k.nmesh), dim3(256), 0, s, L.k, red, state, dt_factor,
                     est_out);
  TICK("dt_final");
  return PG_CHECK_LAUNCH();
}

}  // namespace pangu
