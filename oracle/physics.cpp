// Oracle: fluid state, stress-energy tensor, fluxes, wave speeds,
// reconstruction and the electron-heating models.  TEST INFRASTRUCTURE ONLY.
#include "oracle.h"

namespace oracle {

// u^mu, u_mu, b^mu, b_mu, b^2 from primitives (state_calculation.h:19-69).
void calc_state(const real prim[NPRIM], const real gcov[4][4], const real gcon[4][4], State &s) {
  const real alpha = 1.0 / std::sqrt(-gcon[0][0]);
  // square3(u_tilde, gamma_ij) as dot3 (tensor_algebra.h:19-31): nine terms
  // (v_a * g_ab) * v_b summed left to right.
  const real ut[3] = {prim[UX1], prim[UX2], prim[UX3]};
  real sq = ut[0] * gcov[1][1] * ut[0];
  sq = sq + ut[0] * gcov[1][2] * ut[1];
  sq = sq + ut[0] * gcov[1][3] * ut[2];
  sq = sq + ut[1] * gcov[2][1] * ut[0];
  sq = sq + ut[1] * gcov[2][2] * ut[1];
  sq = sq + ut[1] * gcov[2][3] * ut[2];
  sq = sq + ut[2] * gcov[3][1] * ut[0];
  sq = sq + ut[2] * gcov[3][2] * ut[1];
  sq = sq + ut[2] * gcov[3][3] * ut[2];
  const real lorentz = std::sqrt(1.0 + sq);
  s.ucon[0] = lorentz / alpha;
  s.ucon[1] = prim[UX1] - s.ucon[0] * gcon[0][1] * alpha * alpha;
  s.ucon[2] = prim[UX2] - s.ucon[0] * gcon[0][2] * alpha * alpha;
  s.ucon[3] = prim[UX3] - s.ucon[0] * gcon[0][3] * alpha * alpha;
  if (g_plant == PLANT_SHIFT_SIGN)
    for (int i = 1; i < 4; ++i) s.ucon[i] = prim[UX1 + i - 1] + s.ucon[0] * gcon[0][i] * alpha * alpha;
  for (int r = 0; r < 4; ++r) {
    real acc = 0.0;
    for (int c = 0; c < 4; ++c) acc += gcov[r][c] * s.ucon[c];
    s.ucov[r] = acc;
  }
  s.bcon[0] = prim[BX1] * s.ucov[1] + prim[BX2] * s.ucov[2] + prim[BX3] * s.ucov[3];
  s.bcon[1] = (prim[BX1] + s.bcon[0] * s.ucon[1]) / s.ucon[0];
  s.bcon[2] = (prim[BX2] + s.bcon[0] * s.ucon[2]) / s.ucon[0];
  s.bcon[3] = (prim[BX3] + s.bcon[0] * s.ucon[3]) / s.ucon[0];
  for (int r = 0; r < 4; ++r) {
    real acc = 0.0;
    for (int c = 0; c < 4; ++c) acc += gcov[r][c] * s.bcon[c];
    s.bcov[r] = acc;
  }
  s.bsq = s.bcon[0] * s.bcov[0] + s.bcon[1] * s.bcov[1] + s.bcon[2] * s.bcov[2] +
          s.bcon[3] * s.bcov[3];
}

// T^dir_m (energy_momentum_tensor.h:11-29); recomputes the state as the
// reference does.
void tmunu_row(real gamma, const real prim[NPRIM], const real gcov[4][4],
               const real gcon[4][4], int dir, real t[4]) {
  State s;
  calc_state(prim, gcov, gcon, s);
  const real pg = (gamma - 1.) * prim[ENY];
  const real h = prim[RHO] + gamma * prim[ENY];
  const real etot = h + s.bsq;
  const real ptot = (g_plant == PLANT_EMT_PRESSURE) ? pg : pg + 0.5 * s.bsq;
  for (int m = 0; m < 4; ++m)
    t[m] = etot * s.ucon[dir] * s.ucov[m] + ptot * (real)(dir == m) - s.bcon[dir] * s.bcov[m];
}

// sqrt|g| (rho u^dir, T^dir_0 + rho u^dir, T^dir_i, b^i u^dir - b^dir u^i,
// K rho u^dir, K_e rho u^dir) (contravariant_flux.h:12-39).
void contravariant_flux(real gamma, const real prim[NPRIM], const real gcov[4][4],
                        const real gcon[4][4], real gdet, int dir, real flux[NPRIM]) {
  State s;
  calc_state(prim, gcov, gcon, s);
  real t[4];
  tmunu_row(gamma, prim, gcov, gcon, dir, t);
  flux[RHO] = prim[RHO] * s.ucon[dir];
  flux[ENY] = t[0] + flux[RHO];
  flux[UX1] = t[1];
  flux[UX2] = t[2];
  flux[UX3] = t[3];
  flux[BX1] = s.bcon[1] * s.ucon[dir] - s.bcon[dir] * s.ucon[1];
  flux[BX2] = s.bcon[2] * s.ucon[dir] - s.bcon[dir] * s.ucon[2];
  flux[BX3] = s.bcon[3] * s.ucon[dir] - s.bcon[dir] * s.ucon[3];
  flux[ENT] = prim[ENT] * flux[RHO];
  flux[KEL] = prim[KEL] * flux[RHO];
  const real sg = std::sqrt(std::fabs(gdet));
  for (int n = 0; n < NPRIM; ++n) flux[n] *= sg;
}

static inline real dot4(const real a[4], const real b[4]) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3];
}

// Fast magnetosonic characteristic speeds along dir
// (fast_magnetosonic_speed.h:11-81).
void fast_ms_speed(real gamma, const real prim[NPRIM], const real gcov[4][4],
                   const real gcon[4][4], int dir, real *vmax, real *vmin) {
  real ncon[4] = {0., 0., 0., 0.}, ncov[4] = {0., 0., 0., 0.};
  real tcon[4] = {0., 0., 0., 0.}, tcov[4] = {0., 0., 0., 0.};
  ncov[dir] = 1.0;
  tcov[0] = 1.0;
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) {
      ncon[r] += gcon[r][c] * ncov[c];
      tcon[r] += gcon[r][c] * tcov[c];
    }
  State s;
  calc_state(prim, gcov, gcon, s);
  const real h = prim[RHO] + gamma * prim[ENY];
  const real htot = s.bsq + h;
  const real va2 = s.bsq / htot;
  const real cs2 = gamma * (gamma - 1.) * prim[ENY] / h;
  real cf2 = cs2 + va2 - cs2 * va2;
  if (cf2 < 0.) cf2 = 1e-10;
  if (cf2 > 1.) cf2 = 1.;
  const real nsq = dot4(ncon, ncov);
  const real tsq = dot4(tcon, tcov);
  const real nu = dot4(ncov, s.ucon);
  const real tu = dot4(tcov, s.ucon);
  const real nt = dot4(ncon, tcov);
  const real nu2 = nu * nu;
  const real tu2 = tu * tu;
  const real nutu = nu * tu;
  const real A = tu2 - (tsq + tu2) * cf2;
  const real B = 2. * (nutu - (nt + nutu) * cf2);
  const real C = nu2 - (nsq + nu2) * cf2;
  real disc = B * B - 4. * A * C;
  if ((disc < 0.0) && (disc > -1.e-10)) {
    disc = 0.0;
  } else if (disc < -1.e-10) {
    disc = 0.;
  }
  disc = std::sqrt(disc);
  const real vp = -(-B + disc) / (2. * A);
  const real vm = -(-B - disc) / (2. * A);
  if (vp > vm) {
    *vmax = vp;
    *vmin = vm;
  } else {
    *vmax = vm;
    *vmin = vp;
  }
}

// interpolater_mc.h:16-48.
static inline real minmod3(real a, real b, real c) {
  if (a > 0.0 && b > 0.0 && c > 0.0) return std::fmin(std::fmin(a, b), c);
  if (a < 0.0 && b < 0.0 && c < 0.0) return std::fmax(std::fmax(a, b), c);
  return 0.0;
}
void recon_mc(real qm2, real qm1, real q0, real qp1, real qp2, real *ql_ip1, real *qr_i) {
  (void)qm2;
  (void)qp2;
  const real slope = minmod3(2.0 * (q0 - qm1), 0.5 * (qp1 - qm1), 2.0 * (qp1 - q0));
  *ql_ip1 = q0 + 0.5 * slope;
  *qr_i = q0 - 0.5 * slope;
}

// PPM with Colella-Woodward limiters (interpolater_ppm4.h:29-67).
void recon_ppm4(real qm2, real qm1, real q0, real qp1, real qp2, real *ql_ip1, real *qr_i) {
  real ql = (g_plant == PLANT_PPM_TERM) ? (7.0 * (q0 + qm1) - (qm2)) / 12.0
                                        : (7.0 * (q0 + qm1) - (qm2 + qp1)) / 12.0;
  real qr = (7.0 * (q0 + qp1) - (qm1 + qp2)) / 12.0;
  ql = std::fmax(ql, std::fmin(q0, qm1));
  ql = std::fmin(ql, std::fmax(q0, qm1));
  qr = std::fmax(qr, std::fmin(q0, qp1));
  qr = std::fmin(qr, std::fmax(q0, qp1));
  const real qc = qr - q0;
  const real qd = ql - q0;
  if ((qc * qd) >= 0.0) {
    ql = q0;
    qr = q0;
  } else {
    if (std::fabs(qc) >= 2.0 * std::fabs(qd)) qr = q0 - 2.0 * qd;
    if (std::fabs(qd) >= 2.0 * std::fabs(qc)) ql = q0 - 2.0 * qc;
  }
  *ql_ip1 = qr;
  *qr_i = ql;
}

static inline real clip(real v, real lo, real hi) { return kmax(lo, kmin(v, hi)); }

// heating_model.h:53-109.
real heating_fraction(const Params &p, const HeatCell &c) {
  const real me = 9.1093826e-28;
  const real mp = 1.67262171e-24;
  const real tpr = (p.gamma_p - 1.0) * c.energy / c.rho;
  const real tel = c.adv_e * std::pow(c.rho, p.gamma_e - 1.0);
  switch (p.heat_model) {
    case HM_CONSTANT:
      return p.fel_constant;
    case HM_HOWES: {
      const real trat = tpr / tel;
      const real pres = c.rho * tpr;
      const real beta = kmin(2.0 * pres / c.bsq, 1.0e20);
      const real ltr = std::log(trat) / std::log(10.0);
      const real mbeta = 2.0 - 0.2 * ltr;
      const real c2 = (trat <= 1.0) ? 1.6 / trat : 1.2 / trat;
      const real c3 = (trat <= 1.0) ? 18.0 + 5.0 * ltr : real(18.0);
      const real bp = std::pow(beta, mbeta);
      const real mrat = (g_plant == PLANT_HEAT_HOWES) ? me / mp : mp / me;
      const real qrat = 0.92 * (c2 * c2 + bp) / (c3 * c3 + bp) * std::exp(-1.0 / beta) *
                        std::sqrt(mrat * trat);
      return 1.0 / (1.0 + qrat);
    }
    case HM_KAWAZURA: {
      const real trat = tpr / tel;
      const real pres = c.rho * tpr;
      const real beta = kmin(2.0 * pres / c.bsq, 1.0e20);
      const real ex = (g_plant == PLANT_HEAT_KAWAZURA) ? 1.4 : -1.4;
      const real qiqe = 35.0 / (1.0 + std::pow(beta / 15.0, ex) * std::exp(-0.1 / trat));
      return 1.0 / (1.0 + qiqe);
    }
    case HM_WERNER: {
      const real sigma = c.bsq / c.rho;
      const real sd = (g_plant == PLANT_HEAT_WERNER) ? 2.0 : 5.0;
      return 0.25 * (1.0 + std::sqrt((sigma / sd) / (2.0 + sigma / sd)));
    }
    case HM_ROWAN: {
      const real pres = (p.gamma_p - 1.0) * c.energy;
      const real pg = (p.gamma - 1.0) * c.energy;
      const real beta = 2.0 * pres / c.bsq;
      const real sigma = c.bsq / (c.rho + c.energy + pg);
      const real betamax = ((g_plant == PLANT_HEAT_ROWAN) ? 0.5 : 0.25) / sigma;
      return 0.5 * std::exp(-std::pow(1.0 - beta / betamax, 3.3) /
                            (1.0 + 1.2 * std::pow(sigma, 0.7)));
    }
    case HM_SHARMA: {
      const real trinv = (g_plant == PLANT_HEAT_SHARMA) ? tpr / tel : tel / tpr;
      const real qeqi = 0.33 * std::sqrt(trinv);
      return 1.0 / (1.0 + 1.0 / qeqi);
    }
  }
  return p.fel_constant;
}

// heating_model.h:141-170.
real heating_apply(const Params &p, const HeatCell &c) {
  const real fel = heating_fraction(p, c);
  real diss = (p.gamma_e - 1.0) / (p.gamma - 1.0) * std::pow(c.rho, p.gamma - p.gamma_e) *
              (c.rec_total - c.adv_total);
  if (p.suppress_highb_heat && (c.bsq / c.rho > 1.0)) diss = 0.0;
  if (p.enforce_positive_dissipation) diss = kmax(diss, 0.0);
  const real kel = c.adv_e + fel * diss;
  if (!p.limit_kel && p.heat_model == HM_CONSTANT) return kel;
  const real kmx = c.adv_total * std::pow(c.rho, p.gamma - p.gamma_e) /
                   (p.ratio_min * (p.gamma - 1.0) / (p.gamma_p - 1.0) +
                    (p.gamma - 1.0) / (p.gamma_e - 1.0));
  const real kmn = c.adv_total * std::pow(c.rho, p.gamma - p.gamma_e) /
                   (p.ratio_max * (p.gamma - 1.0) / (p.gamma_p - 1.0) +
                    (p.gamma - 1.0) / (p.gamma_e - 1.0));
  return clip(kel, kmn, kmx);
}

// heating_model.h:172-179.
real clamp_by_ratio(real rmin, real rmax, real total, real electron) {
  const real lo = total / (1.0 + rmax);
  const real hi = total / (1.0 + rmin);
  const real b = kmax(0.0, kmin(electron, total));
  return kmax(lo, kmin(b, hi));
}

}  // namespace oracle
