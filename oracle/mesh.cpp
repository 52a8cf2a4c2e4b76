// Oracle: meshblocks, stage containers and the per-stage task sequence of
// pangu/src/task_list/ideal_grmhd.cc, plus the Parthenon pieces it calls
// (FluxDivergence, UpdateIndependentData, ghost exchange, physical BCs, dt).
// TEST INFRASTRUCTURE ONLY.
#include <algorithm>
#include <cfloat>
#include <cstring>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "oracle.h"
#include "oracle_api.h"
#include "oracle_sim.h"

namespace oracle {

// ------------------------------------------------------------------ set-up
// Uniform mesh generator (Parthenon/Athena++ ComputeMeshGeneratorX with a
// symmetric interval and UniformMeshGeneratorX; mesh edges are exact).
static double mesh_gen(int64_t idx, int64_t n, double xmin, double xmax) {
  const long noff = idx - n / 2;
  const long noffc = idx - (n + 1) / 2;
  const double x = (double)(noff + noffc) / (2.0 * n);
  return 0.5 * (xmin + xmax) + (x * xmax - x * xmin);
}

double mesh_edge(int64_t idx, int64_t n, double xmin, double xmax) {
  return idx == 0 ? xmin : (idx == n ? xmax : mesh_gen(idx, n, xmin, xmax));
}

void Sim::setup(const oracle_cfg &c) {
  cfg = c;
  p.metric = c.metric;
  p.a = c.a;
  p.h = c.h;
  p.gamma = c.gamma;
  p.riemann = c.riemann;
  p.limiter = c.limiter;
  p.recovery = c.recovery;
  p.enable_B = c.enable_B;
  p.enable_heating = c.enable_heating;
  p.cfl = c.cfl;
  p.density_floor = c.density_floor;
  p.energy_floor = c.energy_floor;
  p.density_floor_pow = c.density_floor_pow;
  p.energy_floor_pow = c.energy_floor_pow;
  p.sigma_max = c.sigma_max;
  p.lorentz_max = c.lorentz_max;
  p.heat_model = c.heat_model;
  p.fel_constant = c.fel_constant;
  p.gamma_e = c.gamma_e;
  p.gamma_p = c.gamma_p;
  p.ratio_min = c.ratio_min;
  p.ratio_max = c.ratio_max;
  p.limit_kel = c.limit_kel;
  p.suppress_highb_heat = c.suppress_highb_heat;
  p.enforce_positive_dissipation = c.enforce_positive_dissipation;
  p.excise = c.excise;
  p.r_excise = c.r_excise;
  p.dexcise = c.dexcise;
  p.pexcise = c.pexcise;

  ndim = (c.nx[2] > 1) ? 3 : (c.nx[1] > 1 ? 2 : 1);
  // packed component order (package_registration.cc:84-96)
  np = 0;
  slot_of[np++] = RHO;
  slot_of[np++] = ENY;
  slot_of[np++] = UX1;
  slot_of[np++] = UX2;
  slot_of[np++] = UX3;
  iBX = -1;
  if (p.enable_B) {
    iBX = np;
    slot_of[np++] = BX1;
    slot_of[np++] = BX2;
    slot_of[np++] = BX3;
  }
  iENT = np;
  slot_of[np++] = ENT;
  iKEL = -1;
  if (p.enable_heating) {
    iKEL = np;
    slot_of[np++] = KEL;
  }
  nc = np;

  // blocks: the root grid, refined by the static refinement regions
  // (refine.cpp); without regions, root blocks in x1-fastest order
  std::vector<std::array<int64_t, 4>> leaves;
  make_leaves(leaves);
  nblk = (int)leaves.size();
  blk.resize(nblk);
  leaf_of.clear();
  refined = false;
  int64_t off = 0;
  for (int b = 0; b < nblk; ++b) {
    Block &B = blk[b];
    B.level = (int)leaves[b][0];
    if (B.level > 0) refined = true;
    leaf_of[leaves[b]] = b;
    for (int d = 0; d < 3; ++d) {
      B.lx[d] = (int)leaves[b][1 + d];
      const int64_t nbd = (c.nx[d] > 1) ? ((int64_t)c.nb[d] << B.level) : 1;
      B.nx[d] = c.nx[d] / c.nb[d];
      B.ng[d] = (c.nx[d] > 1) ? c.ng : 0;
      B.n[d] = B.nx[d] + 2 * B.ng[d];
      B.s[d] = B.ng[d];
      B.e[d] = B.ng[d] + B.nx[d] - 1;
      B.xmin[d] = mesh_edge(B.lx[d], nbd, c.xmin[d], c.xmax[d]);
      B.xmax[d] = mesh_edge(B.lx[d] + 1, nbd, c.xmin[d], c.xmax[d]);
      B.dx[d] = (B.xmax[d] - B.xmin[d]) / B.nx[d];
      B.xg[d] = B.xmin[d] - B.ng[d] * B.dx[d];
      B.gofs[d] = B.lx[d] * B.nx[d];
    }
    B.cells = (int64_t)B.n[0] * B.n[1] * B.n[2];
    B.off = off;
    off += B.cells;
  }
  total_cells = off;
  for (int k = 0; k < 2; ++k) {
    P[k].assign(total_cells * np, 0.0);
    U[k].assign(total_cells * nc, 0.0);
  }
  dudt.assign(total_cells * nc, 0.0);
  for (int d = 0; d < 3; ++d) F[d].assign(total_cells * nc, 0.0);
  alfven.assign(total_cells * 3, 0.0);
  emf.assign(total_cells * 3, 0.0);
  flag.assign(total_cells, 0.0);
  flag_rec.assign(total_cells, 0.0);
  mask.assign(total_cells, 0);
  c2p_status.assign(total_cells, 0);
  c2p_iters.assign(total_cells, 0);
  stage1_ready = false;
  time = 0.0;
  dt = c.dt_init;
  dt_est = DBL_MAX;
  ncycle = 0;
#ifdef _OPENMP
  if (c.nthreads > 0) omp_set_num_threads(c.nthreads);
#endif
  build_geometry_memo();
}

// --------------------------------------------------------------- geometry
void Sim::build_geometry_memo() {
  memo = (cfg.cache_geometry != 0) && (p.metric != MET_CKS);
  if (!memo) return;
  gm.resize(nblk);
  for (int b = 0; b < nblk; ++b) {
    const Block &B = blk[b];
    BlockMemo &M = gm[b];
    const int ni = B.n[0], nj = B.n[1];
    M.ni = ni;
    M.nj = nj;
    M.cen.resize((size_t)ni * nj);
    M.f1.resize((size_t)(ni + 1) * nj);
    M.f2.resize((size_t)ni * (nj + 1));
    M.conn.resize((size_t)ni * nj * 64);
    const double x3 = Xc(B, 2, B.s[2]);
#pragma omp parallel for schedule(dynamic, 2)
    for (int j = 0; j <= nj; ++j) {
      for (int i = 0; i <= ni; ++i) {
        if (j < nj && i < ni) {
          const real x[4] = {0.0, Xc(B, 0, i), Xc(B, 1, j), x3};
          Geo &g = M.cen[(size_t)j * ni + i];
          g.ok = metric_at(p.metric, x, p.a, p.h, g.gcov, g.gcon, &g.gdet);
          real cn[4][4][4];
          christoffel(p.metric, x, p.a, p.h, cn);
          std::memcpy(&M.conn[((size_t)j * ni + i) * 64], cn, sizeof(cn));
        }
        if (j < nj) {
          const real x[4] = {0.0, Xf(B, 0, i), Xc(B, 1, j), x3};
          Geo &g = M.f1[(size_t)j * (ni + 1) + i];
          g.ok = metric_at(p.metric, x, p.a, p.h, g.gcov, g.gcon, &g.gdet);
        }
        if (i < ni) {
          const real x[4] = {0.0, Xc(B, 0, i), Xf(B, 1, j), x3};
          Geo &g = M.f2[(size_t)j * ni + i];
          g.ok = metric_at(p.metric, x, p.a, p.h, g.gcov, g.gcon, &g.gdet);
        }
      }
    }
  }
}

// Metric at the centre (loc 0) or the lower face in direction loc (1..3) of
// cell (k,j,i); returns invert4's status.
int Sim::metric(int b, int loc, int k, int j, int i, real gcov[4][4], real gcon[4][4],
                real *gdet) const {
  const Block &B = blk[b];
  if (memo) {
    const BlockMemo &M = gm[b];
    const Geo *g;
    if (loc == 1)
      g = &M.f1[(size_t)j * (M.ni + 1) + i];
    else if (loc == 2)
      g = &M.f2[(size_t)j * M.ni + i];
    else
      g = &M.cen[(size_t)j * M.ni + i];
    std::memcpy(gcov, g->gcov, sizeof(g->gcov));
    if (g->ok) std::memcpy(gcon, g->gcon, sizeof(g->gcon));
    *gdet = g->gdet;
    return g->ok;
  }
  real x[4] = {0.0, Xc(B, 0, i), Xc(B, 1, j), Xc(B, 2, k)};
  if (loc >= 1) x[loc] = Xf(B, loc - 1, loc == 1 ? i : (loc == 2 ? j : k));
  return metric_at(p.metric, x, p.a, p.h, gcov, gcon, gdet);
}

void Sim::connection(int b, int k, int j, int i, real conn[4][4][4]) const {
  const Block &B = blk[b];
  if (memo) {
    std::memcpy(conn, &gm[b].conn[((size_t)j * gm[b].ni + i) * 64], 64 * sizeof(double));
    return;
  }
  const real x[4] = {0.0, Xc(B, 0, i), Xc(B, 1, j), Xc(B, 2, k)};
  christoffel(p.metric, x, p.a, p.h, conn);
}

template <class Fn>
static void for_bkj(int nb, int k0, int k1, int j0, int j1, Fn fn) {
  const long nk = k1 - k0 + 1, nj = j1 - j0 + 1;
  if (nk <= 0 || nj <= 0) return;
  const long tot = (long)nb * nk * nj;
#pragma omp parallel for schedule(dynamic, 4)
  for (long t = 0; t < tot; ++t) {
    const int b = (int)(t / (nk * nj));
    const long r = t % (nk * nj);
    fn(b, (int)(k0 + r / nj), (int)(j0 + r % nj));
  }
}

void Sim::load_prim(int c, int b, int64_t id, real pc[NPRIM]) const {
  for (int s = 0; s < NPRIM; ++s) pc[s] = 0.0;
  for (int q = 0; q < np; ++q) pc[slot_of[q]] = Pc(c, b, q)[id];
}

// ------------------------------------------------------------------ tasks
// FixPrimitive over the entire block (primitive_fixer.cc:15-169).
void Sim::fix_primitive(int c) {
  const Block &B0 = blk[0];
  const double umax2 = p.lorentz_max * p.lorentz_max - 1.0;
  const double e_exc = p.pexcise / (p.gamma - 1.0);
  const double a2 = p.a * p.a;
  for_bkj(nblk, 0, B0.n[2] - 1, 0, B0.n[1] - 1, [&](int b, int k, int j) {
    const Block &B = blk[b];
    double *prho = Pc(c, b, 0), *peny = Pc(c, b, 1), *pu1 = Pc(c, b, 2), *pu2 = Pc(c, b, 3),
           *pu3 = Pc(c, b, 4), *pent = Pc(c, b, iENT);
    double *pkel = (iKEL >= 0) ? Pc(c, b, iKEL) : nullptr;
    for (int i = 0; i < B.n[0]; ++i) {
      const int64_t id = B.idx(k, j, i);
      const real small = 1.0e-20;
      const real sms = kmax(p.sigma_max, small);
      const real rho_before = kmax(prho[id], small);
      const real x = Xc(B, 0, i), y = Xc(B, 1, j), z = Xc(B, 2, k);
      const real xc[4] = {0.0, x, y, z};
      real r = 1.0;
      if (p.metric == MET_MKS) {
        r = std::exp(x);
      } else if (p.metric == MET_CKS) {
        const real rad2 = x * x + y * y + z * z;
        const real sa = (rad2 - a2) * (rad2 - a2) + 4.0 * a2 * z * z;
        r = std::sqrt(0.5 * (rad2 - a2 + std::sqrt(sa)));
      }
      (void)xc;
      real rho_floor = p.density_floor * std::pow(std::fmax(r, 1.0), p.density_floor_pow);
      const real eng_floor = p.energy_floor * std::pow(std::fmax(r, 1.0), p.energy_floor_pow);
      real gcov[4][4], gcon[4][4], gdet;
      metric(b, 0, k, j, i, gcov, gcon, &gdet);
      const real rho = prho[id];
      const real ug = peny[id];
      real pc[NPRIM];
      load_prim(c, b, id, pc);
      State st;
      calc_state(pc, gcov, gcon, st);
      uint8_t m = 0;
      if (rho_floor < ug / sms) rho_floor = ug / sms;
      const real sigma = st.bsq / rho;
      if (sigma > p.sigma_max) {
        rho_floor = st.bsq / sms;
        m |= FL_SIGMA;
      }
      if (prho[id] < rho_floor) { prho[id] = rho_floor; m |= FL_RHO; }
      if (peny[id] < eng_floor) { peny[id] = eng_floor; m |= FL_ENY; }
      const real u1 = pu1[id], u2 = pu2[id], u3 = pu3[id];
      const real uvsq = gcov[1][1] * u1 * u1 + gcov[2][2] * u2 * u2 + gcov[3][3] * u3 * u3 +
                        2.0 * (gcov[1][2] * u1 * u2 + gcov[1][3] * u1 * u3 + gcov[2][3] * u2 * u3);
      if (umax2 > 0.0 && uvsq > umax2) {
        const real f = std::sqrt(umax2 / kmax(uvsq, small));
        pu1[id] *= f;
        pu2[id] *= f;
        pu3[id] *= f;
        m |= FL_GAMMA;
      }
      const real rho_after = kmax(prho[id], small);
      if (pkel && rho_after != rho_before) pkel[id] *= std::pow(rho_after / rho_before, -p.gamma);
      pent[id] = (p.gamma - 1.0) * peny[id] * std::pow(prho[id], -p.gamma);
      if (pkel) pkel[id] = clamp_by_ratio(p.ratio_min, p.ratio_max, pent[id], pkel[id]);
      if (p.metric == MET_CKS && r < p.r_excise) {
        prho[id] = p.dexcise;
        peny[id] = e_exc;
        pu1[id] = 0.0;
        pu2[id] = 0.0;
        pu3[id] = 0.0;
        pent[id] = (p.gamma - 1.0) * e_exc * std::pow(p.dexcise, -p.gamma);
        if (pkel) pkel[id] = e_exc;
        m |= FL_EXCISE;
      }
      mask[B.off + id] = m;
    }
  });
}

// CalculateConservative over the interior (conservative.cc:15-110).
void Sim::calc_conservative(int c) {
  const Block &B0 = blk[0];
  for_bkj(nblk, B0.s[2], B0.e[2], B0.s[1], B0.e[1], [&](int b, int k, int j) {
    const Block &B = blk[b];
    for (int i = B.s[0]; i <= B.e[0]; ++i) {
      const int64_t id = B.idx(k, j, i);
      real gcov[4][4], gcon[4][4], gdet;
      metric(b, 0, k, j, i, gcov, gcon, &gdet);
      real pc[NPRIM], fl[NPRIM];
      load_prim(c, b, id, pc);
      contravariant_flux(p.gamma, pc, gcov, gcon, gdet, 0, fl);
      for (int q = 0; q < nc; ++q) Uc(c, b, q)[id] = fl[slot_of[q]];
    }
  });
}

// One direction of CalculateLAXF / CalculateHLL (laxf.cc:78-393,
// hll.cc:78-398).  Faces are the lower faces of cell index (k,j,i) along dir.
void Sim::sweep(int c, int dir) {
  const Block &B0 = blk[0];
  const int o1 = (B0.n[0] > 1), o2 = (B0.n[1] > 1), o3 = (B0.n[2] > 1);
  int lo[3], hi[3];
  lo[0] = B0.s[0] - o1; hi[0] = B0.e[0] + o1;
  lo[1] = B0.s[1] - o2; hi[1] = B0.e[1] + o2;
  lo[2] = B0.s[2] - o3; hi[2] = B0.e[2] + o3;
  const int d = dir - 1;
  lo[d] = B0.s[d];
  hi[d] = B0.e[d] + 1;
  if (dir == 3 && p.riemann == RS_HLL) {  // hll.cc:295-296: no transverse extension
    lo[0] = B0.s[0]; hi[0] = B0.e[0];
    lo[1] = B0.s[1]; hi[1] = B0.e[1];
  }
  for_bkj(nblk, lo[2], hi[2], lo[1], hi[1], [&](int b, int k, int j) {
    const Block &B = blk[b];
    const int64_t stride = (dir == 1) ? 1 : (dir == 2 ? (int64_t)B.n[0] : (int64_t)B.n[0] * B.n[1]);
    for (int i = lo[0]; i <= hi[0]; ++i) {
      const int64_t id = B.idx(k, j, i);
      real pl[NPRIM], pr[NPRIM];
      for (int s = 0; s < NPRIM; ++s) { pl[s] = 0.0; pr[s] = 0.0; }
      for (int q = 0; q < np; ++q) {
        const double *a = Pc(c, b, q);
        const real qm3 = a[id - 3 * stride], qm2 = a[id - 2 * stride], qm1 = a[id - stride];
        const real q0 = a[id], qp1 = a[id + stride], qp2 = a[id + 2 * stride];
        real ql_i, qr_im1, ql_ip1, qr_i;
        if (p.limiter == LIM_MC) {
          recon_mc(qm3, qm2, qm1, q0, qp1, &ql_i, &qr_im1);
          recon_mc(qm2, qm1, q0, qp1, qp2, &ql_ip1, &qr_i);
        } else {
          recon_ppm4(qm3, qm2, qm1, q0, qp1, &ql_i, &qr_im1);
          recon_ppm4(qm2, qm1, q0, qp1, qp2, &ql_ip1, &qr_i);
        }
        pl[slot_of[q]] = ql_i;
        pr[slot_of[q]] = qr_i;
      }
      real gcov[4][4], gcon[4][4], gdet;
      const int ok = metric(b, dir, k, j, i, gcov, gcon, &gdet);
      if (!ok) {
        // Singular face metric (the polar axis, theta = 0): the reference
        // reads an uninitialised inverse here; this restatement defines the
        // face flux and wave speed as zero.
        alfven[(B.off + 0) * 3 + (int64_t)d * B.cells + id] = 0.0;
        for (int q = 0; q < nc; ++q) Fc(d, b, q)[id] = 0.0;
        continue;
      }
      real vmaxL, vminL, vmaxR, vminR;
      fast_ms_speed(p.gamma, pl, gcov, gcon, dir, &vmaxL, &vminL);
      fast_ms_speed(p.gamma, pr, gcov, gcon, dir, &vmaxR, &vminR);
      const real cmax = std::fabs(kmax(kmax(0., vmaxL), vmaxR));
      const real cmin = std::fabs(kmax(kmax(0., -vminL), -vminR));
      const real cc = kmax(cmax, cmin);
      alfven[(B.off + 0) * 3 + (int64_t)d * B.cells + id] = cc;
      real UL[NPRIM], UR[NPRIM], FL[NPRIM], FR[NPRIM];
      contravariant_flux(p.gamma, pl, gcov, gcon, gdet, 0, UL);
      contravariant_flux(p.gamma, pr, gcov, gcon, gdet, 0, UR);
      contravariant_flux(p.gamma, pl, gcov, gcon, gdet, dir, FL);
      contravariant_flux(p.gamma, pr, gcov, gcon, gdet, dir, FR);
      for (int q = 0; q < nc; ++q) {
        const int s = slot_of[q];
        real f;
        if (p.riemann == RS_LAXF)
          f = 0.5 * (FL[s] + FR[s] - cc * (UR[s] - UL[s]));
        else
          f = (cmax * FL[s] + cmin * FR[s] - cmax * cmin * (UR[s] - UL[s])) / (cmax + cmin + 1e-20);
        Fc(d, b, q)[id] = f;
      }
    }
  });
}

void Sim::fluxes(int c) {
  sweep(c, 1);
  if (ndim >= 2) sweep(c, 2);
  if (ndim == 3) sweep(c, 3);
}

// ConstraintedTransport (constrained_transport.cc:13-130).
void Sim::ct() {
  // The task exists only with a magnetic field (ideal_grmhd.cc:67-68).
  if (!p.enable_B) return;
  const Block &B0 = blk[0];
  const int o1 = (B0.n[0] > 1), o2 = (B0.n[1] > 1), o3 = (B0.n[2] > 1);
  const int E1 = o2 && o3, E2 = o3 && o1, E3 = o1 && o2;
  const int iB1 = iBX, iB2 = iBX + 1, iB3 = iBX + 2;
  for_bkj(nblk, B0.s[2], B0.e[2] + o3, B0.s[1], B0.e[1] + o2, [&](int b, int k, int j) {
    const Block &B = blk[b];
    const int64_t sj = B.n[0], sk = (int64_t)B.n[0] * B.n[1];
    double *e1 = Ec(b, 0), *e2 = Ec(b, 1), *e3 = Ec(b, 2);
    for (int i = B.s[0]; i <= B.e[0] + o1; ++i) {
      const int64_t id = B.idx(k, j, i);
      e1[id] = 0.0;
      e2[id] = 0.0;
      e3[id] = 0.0;
      if (E1)
        e1[id] = 0.25 * (((real)Fc(1, b, iB3)[id] + Fc(1, b, iB3)[id - sk]) -
                         ((real)Fc(2, b, iB2)[id] + Fc(2, b, iB2)[id - sj]));
      if (E2)
        e2[id] = 0.25 * (((real)Fc(2, b, iB1)[id] + Fc(2, b, iB1)[id - 1]) -
                         ((real)Fc(0, b, iB3)[id] + Fc(0, b, iB3)[id - sk]));
      if (E3)
        e3[id] = 0.25 * (((real)Fc(0, b, iB2)[id] + Fc(0, b, iB2)[id - sj]) -
                         ((real)Fc(1, b, iB1)[id] + Fc(1, b, iB1)[id - 1]));
    }
  });
  for_bkj(nblk, B0.s[2], B0.e[2], B0.s[1], B0.e[1], [&](int b, int k, int j) {
    const Block &B = blk[b];
    const int64_t sj = B.n[0], sk = (int64_t)B.n[0] * B.n[1];
    const double *e2 = Ec(b, 1), *e3 = Ec(b, 2);
    for (int i = B.s[0]; i <= B.e[0] + o1; ++i) {
      const int64_t id = B.idx(k, j, i);
      if (o1) Fc(0, b, iB1)[id] = 0;
      if (E2) Fc(0, b, iB3)[id] = -0.5 * ((real)e2[id] + e2[id + sk]);
      if (E3 && g_plant == PLANT_CT_SIGN) Fc(0, b, iB2)[id] = -0.5 * ((real)e3[id] + e3[id + sj]);
      else
      if (E3) Fc(0, b, iB2)[id] = 0.5 * ((real)e3[id] + e3[id + sj]);
    }
  });
  for_bkj(nblk, B0.s[2], B0.e[2], B0.s[1], B0.e[1] + o2, [&](int b, int k, int j) {
    const Block &B = blk[b];
    const int64_t sk = (int64_t)B.n[0] * B.n[1];
    const double *e1 = Ec(b, 0), *e3 = Ec(b, 2);
    for (int i = B.s[0]; i <= B.e[0]; ++i) {
      const int64_t id = B.idx(k, j, i);
      if (E1) Fc(1, b, iB3)[id] = 0.5 * ((real)e1[id] + e1[id + sk]);
      if (o2) Fc(1, b, iB2)[id] = 0.;
      if (E3) Fc(1, b, iB1)[id] = -0.5 * ((real)e3[id] + e3[id + 1]);
    }
  });
  for_bkj(nblk, B0.s[2], B0.e[2] + o3, B0.s[1], B0.e[1], [&](int b, int k, int j) {
    const Block &B = blk[b];
    const int64_t sj = B.n[0];
    const double *e1 = Ec(b, 0), *e2 = Ec(b, 1);
    for (int i = B.s[0]; i <= B.e[0]; ++i) {
      const int64_t id = B.idx(k, j, i);
      if (E1) Fc(2, b, iB2)[id] = -0.5 * ((real)e1[id] + e1[id + sj]);
      if (E2) Fc(2, b, iB1)[id] = 0.5 * ((real)e2[id] + e2[id + 1]);
      if (o3) Fc(2, b, iB3)[id] = 0;
    }
  });
}

// Parthenon FluxDivergence on a uniform Cartesian grid (ideal_grmhd.cc:73-74):
// dU/dt = -[sum_d (A_d F_d(+) - A_d F_d(-))] / V with A_1 = dx2 dx3,
// A_2 = dx1 dx3, A_3 = dx1 dx2, V = dx1 dx2 dx3.
void Sim::flux_div() {
  const Block &B0 = blk[0];
  for_bkj(nblk, B0.s[2], B0.e[2], B0.s[1], B0.e[1], [&](int b, int k, int j) {
    const Block &B = blk[b];
    const int64_t sj = B.n[0], sk = (int64_t)B.n[0] * B.n[1];
    const real a1 = B.dx[1] * B.dx[2], a2 = B.dx[0] * B.dx[2], a3 = B.dx[0] * B.dx[1];
    const real vol = B.dx[0] * B.dx[1] * B.dx[2];
    for (int q = 0; q < nc; ++q) {
      const double *f1 = Fc(0, b, q), *f2 = Fc(1, b, q), *f3 = Fc(2, b, q);
      double *du = dudt.data() + B.off * nc + (int64_t)q * B.cells;
      for (int i = B.s[0]; i <= B.e[0]; ++i) {
        const int64_t id = B.idx(k, j, i);
        real s = (a1 * f1[id + 1] - a1 * f1[id]);
        if (ndim >= 2) s += (a2 * f2[id + sj] - a2 * f2[id]);
        if (ndim == 3) s += (a3 * f3[id + sk] - a3 * f3[id]);
        du[id] = -s / vol;
      }
    }
  });
}

// UpdateIndependentData: U(out) = U(base) + beta dt dU/dt (ideal_grmhd.cc:75-76).
void Sim::update(double bdt, int out) {
  const Block &B0 = blk[0];
  for_bkj(nblk, B0.s[2], B0.e[2], B0.s[1], B0.e[1], [&](int b, int k, int j) {
    const Block &B = blk[b];
    for (int q = 0; q < nc; ++q) {
      const double *in = Uc(0, b, q);
      const double *du = dudt.data() + B.off * nc + (int64_t)q * B.cells;
      double *o = Uc(out, b, q);
      for (int i = B.s[0]; i <= B.e[0]; ++i) {
        const int64_t id = B.idx(k, j, i);
        o[id] = in[id] + (real)bdt * du[id];
      }
    }
  });
}

// AddGeometricSource (source_term.cc:15-121).
void Sim::source(int c, double bdt) {
  const Block &B0 = blk[0];
  for_bkj(nblk, B0.s[2], B0.e[2], B0.s[1], B0.e[1], [&](int b, int k, int j) {
    const Block &B = blk[b];
    for (int i = B.s[0]; i <= B.e[0]; ++i) {
      const int64_t id = B.idx(k, j, i);
      real gcov[4][4], gcon[4][4], gdet;
      metric(b, 0, k, j, i, gcov, gcon, &gdet);
      const real sg = std::sqrt(std::fabs(gdet));
      real conn[4][4][4];
      connection(b, k, j, i, conn);
      real pc[NPRIM];
      load_prim(c, b, id, pc);
      real T[4][4];
      for (int r = 0; r < 4; ++r) tmunu_row(p.gamma, pc, gcov, gcon, r, T[r]);
      for (int dd = 0; dd < 4; ++dd) {
        real con = 0.0;
        for (int r = 0; r < 4; ++r)
          for (int cc = 0; cc < 4; ++cc) con += T[r][cc] * conn[cc][dd][r];
        double *u = Uc(c, b, dd == 0 ? 1 : 2 + (dd - 1));
        u[id] += con * sg * bdt;
      }
    }
  });
}

// Recovery (invertor.cc:18-210).
void Sim::recovery(int c) {
  const Block &B0 = blk[0];
  const bool chainer = (p.recovery == REC_CHAINER) || (p.recovery == REC_1DVSQ);
  const double a2 = p.a * p.a;
  for_bkj(nblk, B0.s[2], B0.e[2], B0.s[1], B0.e[1], [&](int b, int k, int j) {
    const Block &B = blk[b];
    for (int i = B.s[0]; i <= B.e[0]; ++i) {
      const int64_t id = B.idx(k, j, i);
      const int64_t gid = B.off + id;
      if (p.metric == MET_CKS) {
        const real x = Xc(B, 0, i), y = Xc(B, 1, j), z = Xc(B, 2, k);
        const real rad2 = x * x + y * y + z * z;
        const real sa = (rad2 - a2) * (rad2 - a2) + 4.0 * a2 * z * z;
        const real rbl = std::sqrt(0.5 * (rad2 - a2 + std::sqrt(sa)));
        if (rbl < p.r_excise) {
          if (chainer) flag[gid] = 0;
          c2p_status[gid] = -1;
          c2p_iters[gid] = 0;
          continue;
        }
      }
      real gcov[4][4], gcon[4][4], gdet;
      metric(b, 0, k, j, i, gcov, gcon, &gdet);
      if (chainer) flag[gid] = 0;
      const real sg = std::sqrt(std::fabs(gdet));
      const real alpha = 1.0 / std::sqrt(-gcon[0][0]);
      real Uc9[NPRIM_RECV], Pc9[NPRIM_RECV];
      for (int s = 0; s < NPRIM_RECV; ++s) { Uc9[s] = 0.0; Pc9[s] = 0.0; }
      {
        int q = 0;
        Uc9[RHO] = Uc(c, b, q++)[id];
        Uc9[ENY] = Uc(c, b, q++)[id];
        Uc9[UX1] = Uc(c, b, q++)[id];
        Uc9[UX2] = Uc(c, b, q++)[id];
        Uc9[UX3] = Uc(c, b, q++)[id];
        if (p.enable_B) {
          Uc9[BX1] = Uc(c, b, q++)[id];
          Uc9[BX2] = Uc(c, b, q++)[id];
          Uc9[BX3] = Uc(c, b, q++)[id];
        }
        Uc9[ENT] = Uc(c, b, q++)[id];
        q = 0;
        Pc9[RHO] = Pc(c, b, q++)[id];
        Pc9[ENY] = Pc(c, b, q++)[id];
        Pc9[UX1] = Pc(c, b, q++)[id];
        Pc9[UX2] = Pc(c, b, q++)[id];
        Pc9[UX3] = Pc(c, b, q++)[id];
        if (p.enable_B) {
          Pc9[BX1] = Pc(c, b, q++)[id];
          Pc9[BX2] = Pc(c, b, q++)[id];
          Pc9[BX3] = Pc(c, b, q++)[id];
        }
        Pc9[ENT] = Pc(c, b, q++)[id];
      }
      const real isg = 1.0 / sg;
      const real aos = alpha * isg;
      if (p.enable_B) {
        Pc(c, b, iBX)[id] = Uc9[BX1] * isg;
        Pc(c, b, iBX + 1)[id] = Uc9[BX2] * isg;
        Pc(c, b, iBX + 2)[id] = Uc9[BX3] * isg;
      }
      real Uh[NPRIM_RECV], Ph[NPRIM_RECV];
      Uh[RHO] = aos * Uc9[RHO];
      Uh[ENY] = aos * (Uc9[ENY] - Uc9[RHO]);
      Uh[UX1] = aos * Uc9[UX1];
      Uh[UX2] = aos * Uc9[UX2];
      Uh[UX3] = aos * Uc9[UX3];
      Uh[BX1] = aos * Uc9[BX1];
      Uh[BX2] = aos * Uc9[BX2];
      Uh[BX3] = aos * Uc9[BX3];
      Uh[ENT] = aos * Uc9[ENT];
      Ph[RHO] = Pc9[RHO];
      Ph[ENY] = Pc9[ENY];
      Ph[UX1] = Pc9[UX1];
      Ph[UX2] = Pc9[UX2];
      Ph[UX3] = Pc9[UX3];
      Ph[BX1] = 0.0;
      Ph[BX2] = 0.0;
      Ph[BX3] = 0.0;
      if (p.enable_B) {
        Ph[BX1] = alpha * Uc9[BX1] * isg;
        Ph[BX2] = alpha * Uc9[BX2] * isg;
        Ph[BX3] = alpha * Uc9[BX3] * isg;
      }
      Ph[ENT] = Pc9[ENT];
      int st = 0, it = 0, itn = 0;
      if (p.recovery == REC_KASTAUN) {
        st = c2p_kastaun(Uh, Ph, p.gamma, gcov, gcon, &it);
        c2p_status[gid] = st;
        c2p_iters[gid] = it;
        if (st != 0) continue;
      } else if (p.recovery == REC_1DVSQ) {
        st = c2p_1dvsq(Uh, Ph, p.gamma, gcov, gcon, &it);
        flag[gid] = (st == 0) ? 1.0 : 0.0;
        c2p_status[gid] = st;
        c2p_iters[gid] = it;
      } else {
        st = c2p_2d(Uh, Ph, p.gamma, gcov, gcon, &it);
        itn = it;
        if (st != 0) {
          st = c2p_1dvsq(Uh, Ph, p.gamma, gcov, gcon, &it);
          itn += it;
          if (st != 0) {
            st = c2p_1d(Uh, Ph, p.gamma, gcov, gcon, &it);
            itn += it;
          }
        }
        flag[gid] = (st == 0) ? 1.0 : 0.0;
        c2p_status[gid] = st;
        c2p_iters[gid] = itn;
      }
      Pc(c, b, 0)[id] = Ph[RHO];
      Pc(c, b, 1)[id] = Ph[ENY];
      Pc(c, b, 2)[id] = Ph[UX1];
      Pc(c, b, 3)[id] = Ph[UX2];
      Pc(c, b, 4)[id] = Ph[UX3];
    }
  });
  flag_rec = flag;
}

// ApplyElectronHeating (electron_heating.cc:14-112).
void Sim::heating(int c) {
  const Block &B0 = blk[0];
  for_bkj(nblk, B0.s[2], B0.e[2], B0.s[1], B0.e[1], [&](int b, int k, int j) {
    const Block &B = blk[b];
    for (int i = B.s[0]; i <= B.e[0]; ++i) {
      const int64_t id = B.idx(k, j, i);
      real gcov[4][4], gcon[4][4], gdet;
      metric(b, 0, k, j, i, gcov, gcon, &gdet);
      real pc[NPRIM];
      load_prim(c, b, id, pc);
      State st;
      calc_state(pc, gcov, gcon, st);
      const real crho = Uc(c, b, 0)[id];
      HeatCell cell;
      cell.rho = Pc(c, b, 0)[id];
      cell.energy = Pc(c, b, 1)[id];
      cell.bsq = st.bsq;
      cell.adv_total = Uc(c, b, iENT)[id] / crho;
      cell.rec_total = Pc(c, b, iENT)[id];
      cell.adv_e = Uc(c, b, iKEL)[id] / crho;
      Pc(c, b, iKEL)[id] = heating_apply(p, cell);
    }
  });
}

// FixRecovery (recovery_fixer.cc:12-119) with the serial (lexicographic,
// in place) ordering of a single-threaded Kokkos host build.  Planes k are
// independent (the stencil is in i and j only).
void Sim::fix_recovery(int c) {
  const Block &B0 = blk[0];
  const int oj = (B0.s[1] != B0.e[1]) ? 1 : 0;
  const double a2 = p.a * p.a;
  const long nk = B0.e[2] - B0.s[2] + 1;
#pragma omp parallel for schedule(dynamic, 1)
  for (long t = 0; t < (long)nblk * nk; ++t) {
    const int b = (int)(t / nk);
    const int k = B0.s[2] + (int)(t % nk);
    const Block &B = blk[b];
    const int64_t sj = B.n[0];
    double *fl = flag.data() + B.off;
    for (int j = B.s[1]; j <= B.e[1]; ++j)
      for (int i = B.s[0]; i <= B.e[0]; ++i) {
        const int64_t id = B.idx(k, j, i);
        if (fl[id] != 0) continue;
        if (p.metric == MET_CKS) {
          const real x = Xc(B, 0, i), y = Xc(B, 1, j), z = Xc(B, 2, k);
          const real rad2 = x * x + y * y + z * z;
          const real sa = (rad2 - a2) * (rad2 - a2) + 4.0 * a2 * z * z;
          if (std::sqrt(0.5 * (rad2 - a2 + std::sqrt(sa))) < p.r_excise) continue;
        }
        const int64_t jp = oj * sj;
        const int pf1 = (int)fl[id + jp - 1], pf2 = (int)fl[id + jp], pf3 = (int)fl[id + jp + 1];
        const int pf4 = (int)fl[id + 1];
        const int pf5 = (int)fl[id - jp + 1], pf6 = (int)fl[id - jp], pf7 = (int)fl[id - jp - 1];
        const int pf8 = (int)fl[id - 1];
        for (int q = 0; q < np; ++q) {
          double *a = Pc(c, b, q);
          if (pf2 && pf4 && pf6 && pf8)
            a[id] = 0.25 * (a[id + jp] + a[id - jp] + a[id - 1] + a[id + 1]);
          else if (pf1 && pf3 && pf5 && pf7)
            a[id] = 0.25 * (a[id + jp + 1] + a[id + jp - 1] + a[id - jp + 1] + a[id - jp - 1]);
          else
            a[id] = 0.125 * (a[id + jp - 1] + a[id + jp] + a[id + jp + 1] + a[id + 1] +
                             a[id - jp + 1] + a[id - jp] + a[id - jp - 1] + a[id - 1]);
        }
        Pc(c, b, 2)[id] = 0.;
        Pc(c, b, 3)[id] = 0.;
        Pc(c, b, 4)[id] = 0.;
        fl[id] = 1;
      }
  }
}

// Same-level ghost exchange of the fields read in ghost zones (primitives
// and flag) from face, edge and corner neighbours; periodic dims wrap.
void Sim::exchange(int c) {
  if (refined) {
    exchange_refined(c);
    return;
  }
  const int nbd[3] = {cfg.nb[0], cfg.nb[1], cfg.nb[2]};
  for (int b = 0; b < nblk; ++b) {
    const Block &B = blk[b];
#pragma omp parallel for schedule(static)
    for (int k = 0; k < B.n[2]; ++k)
      for (int j = 0; j < B.n[1]; ++j)
        for (int i = 0; i < B.n[0]; ++i) {
          const int ix[3] = {i, j, k};
          int o[3];
          bool interior = true;
          for (int d = 0; d < 3; ++d) {
            o[d] = (ix[d] < B.s[d]) ? -1 : (ix[d] > B.e[d] ? 1 : 0);
            if (o[d] != 0) interior = false;
          }
          if (interior) continue;
          int lx[3];
          bool ok = true;
          for (int d = 0; d < 3; ++d) {
            lx[d] = B.lx[d] + o[d];
            if (lx[d] < 0 || lx[d] >= nbd[d]) {
              if (cfg.bc[d][0] == BC_PERIODIC)
                lx[d] = (lx[d] + nbd[d]) % nbd[d];
              else
                ok = false;
            }
          }
          if (!ok) continue;
          const int nb = lx[0] + nbd[0] * (lx[1] + nbd[1] * lx[2]);
          const Block &N = blk[nb];
          const int si = i - o[0] * B.nx[0], sjj = j - o[1] * B.nx[1], sk = k - o[2] * B.nx[2];
          const int64_t dst = B.idx(k, j, i), src = N.idx(sk, sjj, si);
          for (int q = 0; q < np; ++q) Pc(c, b, q)[dst] = Pc(c, nb, q)[src];
          flag[B.off + dst] = flag[N.off + src];
        }
  }
}

// Physical boundary conditions (outflow copies the edge cell, reflecting
// mirrors and flips the normal component of the vector fields u~ and B),
// applied X1, X2, X3 in turn over the full extent of the other dims.
void Sim::apply_bc(int c) {
  for (int b = 0; b < nblk; ++b) {
    const Block &B = blk[b];
    for (int d = 0; d < ndim; ++d) {
      for (int side = 0; side < 2; ++side) {
        const int kind = cfg.bc[d][side];
        if (kind == BC_PERIODIC) continue;
        if (side == 0 && B.lx[d] != 0) continue;
        if (side == 1 && (int64_t)B.lx[d] != ((int64_t)cfg.nb[d] << B.level) - 1) continue;
        const int g0 = (side == 0) ? 0 : B.e[d] + 1;
        const int g1 = (side == 0) ? B.s[d] - 1 : B.n[d] - 1;
        for (int k = 0; k < B.n[2]; ++k)
          for (int j = 0; j < B.n[1]; ++j)
            for (int i = 0; i < B.n[0]; ++i) {
              int ix[3] = {i, j, k};
              if (ix[d] < g0 || ix[d] > g1) continue;
              int sx[3] = {i, j, k};
              if (kind == BC_OUTFLOW)
                sx[d] = (side == 0) ? B.s[d] : B.e[d];
              else
                sx[d] = (side == 0) ? (2 * B.s[d] - 1 - ix[d]) : (2 * B.e[d] + 1 - ix[d]);
              const int64_t dst = B.idx(k, j, i), src = B.idx(sx[2], sx[1], sx[0]);
              for (int q = 0; q < np; ++q) {
                double v = Pc(c, b, q)[src];
                if (kind == BC_REFLECT) {
                  const int s = slot_of[q];
                  if ((s >= UX1 && s <= UX3 && s - UX1 == d) || (s >= BX1 && s <= BX3 && s - BX1 == d))
                    v = -v;
                }
                Pc(c, b, q)[dst] = v;
              }
              flag[B.off + dst] = flag[B.off + src];
            }
      }
    }
  }
}

// EstimateTimestepMesh (timestep_estimation.cc:13-102): per block the
// harmonic combination of min(cfl dx_d / c_d), then the minimum over blocks.
double Sim::estimate_dt() const {
  double gmin = DBL_MAX;
  for (int b = 0; b < nblk; ++b) {
    const Block &B = blk[b];
    double m[3] = {DBL_MAX, DBL_MAX, DBL_MAX};
    for (int d = 0; d < 3; ++d) {
      if (B.n[d] <= 1) continue;
      const double *al = alfven.data() + B.off * 3 + (int64_t)d * B.cells;
      double mm = DBL_MAX;
      for (int k = B.s[2]; k <= B.e[2]; ++k)
        for (int j = B.s[1]; j <= B.e[1]; ++j)
          for (int i = B.s[0]; i <= B.e[0]; ++i) {
            const real dtc = (real)p.cfl * B.dx[d] / al[B.idx(k, j, i)];
            if (mm > dtc) mm = dtc;
          }
      m[d] = mm;
    }
    const double t = 1.0 / (1.0 / m[0] + 1.0 / m[1] + 1.0 / m[2]);
    if (t < gmin) gmin = t;
  }
  return gmin;
}

// ----------------------------------------------------------------- driver
// One RK stage of Simulator::MakeTaskCollection (ideal_grmhd.cc:23-120).
// rk2: beta = {1, 1/2}, stage_name = {"base", "1", "base"}; the stage uses
// beta[stage % 2] (:29).
#ifdef ORACLE_COUNTING
uint64_t g_ops[OP_N];
uint64_t g_task_ops[ORACLE_NTASKS][OP_N];
// Runs one task and adds the operations it performed to its bucket.
#define COUNTED(task, call)                                                 \
  do {                                                                      \
    uint64_t s_[OP_N];                                                      \
    for (int q_ = 0; q_ < OP_N; ++q_) s_[q_] = g_ops[q_];                   \
    call;                                                                   \
    for (int q_ = 0; q_ < OP_N; ++q_) g_task_ops[task][q_] += g_ops[q_] - s_[q_]; \
  } while (0)
#else
#define COUNTED(task, call) call
#endif

void Sim::run_stage(int stage) {
  const double beta = (stage == 1) ? 0.5 : 1.0;
  const int c0 = (stage == 1) ? 0 : 1;
  const int c1 = (stage == 1) ? 1 : 0;
  ensure_stage_container();
  const double bdt = beta * dt;
  COUNTED(ORT_FIX_PRIMITIVE, fix_primitive(c0));
  COUNTED(ORT_CONSERVATIVE, calc_conservative(c0));
  COUNTED(ORT_FLUXES, fluxes(c0));
  if (p.enable_B) COUNTED(ORT_CT, ct());
  if (refined) COUNTED(ORT_FLUX_CORRECTION, flux_correction());
  COUNTED(ORT_FLUX_DIV, flux_div());
  COUNTED(ORT_UPDATE, update(bdt, c1));
  if (cfg.exchange_after == 0) COUNTED(ORT_EXCHANGE, exchange(c1));
  COUNTED(ORT_SOURCE, source(c1, bdt));
  COUNTED(ORT_RECOVERY, recovery(c1));
  if (p.enable_heating) COUNTED(ORT_HEATING, heating(c1));
  if (p.recovery != REC_KASTAUN) COUNTED(ORT_FIX_RECOVERY, fix_recovery(c1));
  if (cfg.exchange_after == 1) COUNTED(ORT_EXCHANGE, exchange(c1));
  COUNTED(ORT_BC, apply_bc(c1));
  if (stage == 2) COUNTED(ORT_TIMESTEP, dt_est = estimate_dt());
}

void Sim::ensure_stage_container() {
  if (stage1_ready) return;
  if (cfg.stage_init == 1) P[1] = P[0];
  stage1_ready = true;
}

void Sim::cycle() {
  run_stage(1);
  run_stage(2);
  time += dt;
  ncycle += 1;
  dt = kmin(cfg.dt_factor * dt, dt_est);
}

}  // namespace oracle
