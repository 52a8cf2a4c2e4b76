// Device functions of the GRMHD update path (gfx950, FP64).
//
// Each function keeps the reference's operand order so results match an
// FP64 host evaluation bit for bit (builds use -ffp-contract=off; division
// and sqrt are IEEE-rounded on gfx950, and the divisions of exact_div.h
// return the IEEE quotient).  Differences from a host evaluation
// can only come from pow/exp/log, whose device implementations may differ by
// an ulp.  The metric arrives from host-built tables instead of being
// recomputed, and a face's fluid state is computed once per side instead of
// the reference's ten recomputations per face (identical values).
#pragma once
#include <hip/hip_runtime.h>

#include "common.h"
#include "exact_div.h"

namespace pangu {

struct Phys {
  double gamma, cfl;
  double density_floor, energy_floor, density_floor_pow, energy_floor_pow;
  double sigma_max, lorentz_max;
  int heat_model;
  double fel_constant, gamma_e, gamma_p, ratio_min, ratio_max;
  int limit_kel, suppress_highb_heat, enforce_positive_dissipation;
  int metric, excise;
  double r_excise, dexcise, pexcise;
  double a;
};

// Kokkos::max / Kokkos::min semantics.
__device__ __forceinline__ double kmax(double a, double b) { return (a < b) ? b : a; }
__device__ __forceinline__ double kmin(double a, double b) { return (b < a) ? b : a; }

// Fluid state u^mu, u_mu, b^mu, b_mu, b^2 (state_calculation.h:19-69).
// g: gcov row-major; gc0: gcon[0][0..3]; alpha = 1/sqrt(-gcon00).
struct St {
  double ucon[4], ucov[4], bcon[4], bcov[4], bsq;
};

__device__ __forceinline__ void calc_state(const double p[NPRIM], const double g[16],
                                           const double gc0[4], double alpha, St &s) {
  const double v1 = p[UX1], v2 = p[UX2], v3 = p[UX3];
  double sq = v1 * g[5] * v1;
  sq = sq + v1 * g[6] * v2;
  sq = sq + v1 * g[7] * v3;
  sq = sq + v2 * g[9] * v1;
  sq = sq + v2 * g[10] * v2;
  sq = sq + v2 * g[11] * v3;
  sq = sq + v3 * g[13] * v1;
  sq = sq + v3 * g[14] * v2;
  sq = sq + v3 * g[15] * v3;
  const double lor = sqrt(1.0 + sq);
  s.ucon[0] = lor / alpha;
  s.ucon[1] = v1 - s.ucon[0] * gc0[1] * alpha * alpha;
  s.ucon[2] = v2 - s.ucon[0] * gc0[2] * alpha * alpha;
  s.ucon[3] = v3 - s.ucon[0] * gc0[3] * alpha * alpha;
#pragma unroll
  for (int r = 0; r < 4; ++r)
    s.ucov[r] = g[4 * r] * s.ucon[0] + g[4 * r + 1] * s.ucon[1] + g[4 * r + 2] * s.ucon[2] +
                g[4 * r + 3] * s.ucon[3];
  s.bcon[0] = p[BX1] * s.ucov[1] + p[BX2] * s.ucov[2] + p[BX3] * s.ucov[3];
  const double ru0 = 1.0 / s.ucon[0];
  s.bcon[1] = div_by(p[BX1] + s.bcon[0] * s.ucon[1], s.ucon[0], ru0);
  s.bcon[2] = div_by(p[BX2] + s.bcon[0] * s.ucon[2], s.ucon[0], ru0);
  s.bcon[3] = div_by(p[BX3] + s.bcon[0] * s.ucon[3], s.ucon[0], ru0);
#pragma unroll
  for (int r = 0; r < 4; ++r)
    s.bcov[r] = g[4 * r] * s.bcon[0] + g[4 * r + 1] * s.bcon[1] + g[4 * r + 2] * s.bcon[2] +
                g[4 * r + 3] * s.bcon[3];
  s.bsq = s.bcon[0] * s.bcov[0] + s.bcon[1] * s.bcov[1] + s.bcon[2] * s.bcov[2] +
          s.bcon[3] * s.bcov[3];
}

// T^dir_m (energy_momentum_tensor.h:11-29).  The reference's
// p_tot*(dir==m) term is added only where it is non-zero (adding +0 leaves
// the sum unchanged).
template <int DIR>
__device__ __forceinline__ double tmunu(const St &s, double etot, double ptot, int m) {
  double v = etot * s.ucon[DIR] * s.ucov[m];
  if (m == DIR) v = v + ptot;
  return v - s.bcon[DIR] * s.bcov[m];
}

// sqrt|g| * F^dir (contravariant_flux.h:12-39) for every slot of the vector.
template <int DIR>
__device__ __forceinline__ void contravariant_flux(double gamma, const double p[NPRIM], const St &s,
                                                   double sg, double f[NPRIM]) {
  const double pg = (gamma - 1.) * p[ENY];
  const double h = p[RHO] + gamma * p[ENY];
  const double etot = h + s.bsq;
  const double ptot = pg + 0.5 * s.bsq;
  const double mflux = p[RHO] * s.ucon[DIR];
  f[RHO] = mflux * sg;
  f[ENY] = (tmunu<DIR>(s, etot, ptot, 0) + mflux) * sg;
  f[UX1] = tmunu<DIR>(s, etot, ptot, 1) * sg;
  f[UX2] = tmunu<DIR>(s, etot, ptot, 2) * sg;
  f[UX3] = tmunu<DIR>(s, etot, ptot, 3) * sg;
  f[BX1] = (s.bcon[1] * s.ucon[DIR] - s.bcon[DIR] * s.ucon[1]) * sg;
  f[BX2] = (s.bcon[2] * s.ucon[DIR] - s.bcon[DIR] * s.ucon[2]) * sg;
  f[BX3] = (s.bcon[3] * s.ucon[DIR] - s.bcon[DIR] * s.ucon[3]) * sg;
  f[ENT] = (p[ENT] * mflux) * sg;
  f[KEL] = (p[KEL] * mflux) * sg;
}

// Fast magnetosonic speeds along DIR (fast_magnetosonic_speed.h:11-81).  With
// n_mu = e^DIR and t_mu = e^0 the reference's contractions reduce exactly to
// n^2 = gcon[DIR][DIR], t^2 = gcon00, n.t = gcon[0][DIR], n.u = u^DIR, t.u = u^0.
template <int DIR>
__device__ __forceinline__ void fast_ms(double gamma, const double p[NPRIM], const St &s,
                                        double gcon00, double gcon0d, double gcondd,
                                        double &vmax, double &vmin) {
  const double h = p[RHO] + gamma * p[ENY];
  const double htot = s.bsq + h;
  const double va2 = s.bsq / htot;
  const double cs2 = gamma * (gamma - 1.) * p[ENY] / h;
  double cf2 = cs2 + va2 - cs2 * va2;
  if (cf2 < 0.) cf2 = 1e-10;
  if (cf2 > 1.) cf2 = 1.;
  const double nu = s.ucon[DIR], tu = s.ucon[0];
  const double nu2 = nu * nu, tu2 = tu * tu, nutu = nu * tu;
  const double A = tu2 - (gcon00 + tu2) * cf2;
  const double B = 2. * (nutu - (gcon0d + nutu) * cf2);
  const double C = nu2 - (gcondd + nu2) * cf2;
  double disc = B * B - 4. * A * C;
  if ((disc < 0.0) && (disc > -1.e-10)) {
    disc = 0.0;
  } else if (disc < -1.e-10) {
    disc = 0.;
  }
  disc = sqrt(disc);
  const double a2 = 2. * A, ra2 = 1.0 / a2;
  const double vp = div_by(-(-B + disc), a2, ra2);
  const double vm = div_by(-(-B - disc), a2, ra2);
  if (vp > vm) {
    vmax = vp;
    vmin = vm;
  } else {
    vmax = vm;
    vmin = vp;
  }
}

// MC (interpolater_mc.h:16-48): values at the right face of the cell
// (ql_ip1) and its left face (qr_i) from the centre triplet.
__device__ __forceinline__ void recon_mc(double qm1, double q0, double qp1, double &ql_ip1,
                                         double &qr_i) {
  const double a = 2.0 * (q0 - qm1), b = 0.5 * (qp1 - qm1), c = 2.0 * (qp1 - q0);
  double slope = 0.0;
  if (a > 0.0 && b > 0.0 && c > 0.0)
    slope = fmin(fmin(a, b), c);
  else if (a < 0.0 && b < 0.0 && c < 0.0)
    slope = fmax(fmax(a, b), c);
  ql_ip1 = q0 + 0.5 * slope;
  qr_i = q0 - 0.5 * slope;
}

// PPM with Colella-Woodward limiting (interpolater_ppm4.h:29-67).
__device__ __forceinline__ void recon_ppm4(double qm2, double qm1, double q0, double qp1,
                                           double qp2, double &ql_ip1, double &qr_i) {
  double ql = div12(7.0 * (q0 + qm1) - (qm2 + qp1));
  double qr = div12(7.0 * (q0 + qp1) - (qm1 + qp2));
  ql = fmax(ql, fmin(q0, qm1));
  ql = fmin(ql, fmax(q0, qm1));
  qr = fmax(qr, fmin(q0, qp1));
  qr = fmin(qr, fmax(q0, qp1));
  const double qc = qr - q0, qd = ql - q0;
  if ((qc * qd) >= 0.0) {
    ql = q0;
    qr = q0;
  } else {
    if (fabs(qc) >= 2.0 * fabs(qd)) qr = q0 - 2.0 * qd;
    if (fabs(qd) >= 2.0 * fabs(qc)) ql = q0 - 2.0 * qc;
  }
  ql_ip1 = qr;
  qr_i = ql;
}

// ---------------------------------------------------------- electron heating
struct HeatIn {
  double rho, energy, bsq, adv_total, rec_total, adv_e;
};

// heating_model.h:53-109.
__device__ __forceinline__ double heat_fraction(const Phys &P, const HeatIn &c) {
  const double me = 9.1093826e-28, mp = 1.67262171e-24;
  const double tpr = (P.gamma_p - 1.0) * c.energy / c.rho;
  const double tel = c.adv_e * pow(c.rho, P.gamma_e - 1.0);
  switch (P.heat_model) {
    case PANGU_HEAT_CONSTANT: return P.fel_constant;
    case PANGU_HEAT_HOWES: {
      const double trat = tpr / tel;
      const double pres = c.rho * tpr;
      const double beta = kmin(2.0 * pres / c.bsq, 1.0e20);
      const double ltr = log(trat) / log(10.0);
      const double mbeta = 2.0 - 0.2 * ltr;
      const double c2 = (trat <= 1.0) ? 1.6 / trat : 1.2 / trat;
      const double c3 = (trat <= 1.0) ? 18.0 + 5.0 * ltr : 18.0;
      const double bp = pow(beta, mbeta);
      const double qrat =
          0.92 * (c2 * c2 + bp) / (c3 * c3 + bp) * exp(-1.0 / beta) * sqrt(mp / me * trat);
      return 1.0 / (1.0 + qrat);
    }
    case PANGU_HEAT_KAWAZURA: {
      const double trat = tpr / tel;
      const double pres = c.rho * tpr;
      const double beta = kmin(2.0 * pres / c.bsq, 1.0e20);
      const double qiqe = 35.0 / (1.0 + pow(beta / 15.0, -1.4) * exp(-0.1 / trat));
      return 1.0 / (1.0 + qiqe);
    }
    case PANGU_HEAT_WERNER: {
      const double sigma = c.bsq / c.rho;
      return 0.25 * (1.0 + sqrt((sigma / 5.0) / (2.0 + sigma / 5.0)));
    }
    case PANGU_HEAT_ROWAN: {
      const double pres = (P.gamma_p - 1.0) * c.energy;
      const double pg = (P.gamma - 1.0) * c.energy;
      const double beta = 2.0 * pres / c.bsq;
      const double sigma = c.bsq / (c.rho + c.energy + pg);
      const double betamax = 0.25 / sigma;
      return 0.5 * exp(-pow(1.0 - beta / betamax, 3.3) / (1.0 + 1.2 * pow(sigma, 0.7)));
    }
    case PANGU_HEAT_SHARMA: {
      const double trinv = tel / tpr;
      const double qeqi = 0.33 * sqrt(trinv);
      return 1.0 / (1.0 + 1.0 / qeqi);
    }
  }
  return P.fel_constant;
}

// heating_model.h:141-170.
__device__ __forceinline__ double heat_apply(const Phys &P, const HeatIn &c) {
  const double fel = heat_fraction(P, c);
  const double rpow = pow(c.rho, P.gamma - P.gamma_e);
  double diss = (P.gamma_e - 1.0) / (P.gamma - 1.0) * rpow * (c.rec_total - c.adv_total);
  if (P.suppress_highb_heat && (c.bsq / c.rho > 1.0)) diss = 0.0;
  if (P.enforce_positive_dissipation) diss = kmax(diss, 0.0);
  const double kel = c.adv_e + fel * diss;
  if (!P.limit_kel && P.heat_model == PANGU_HEAT_CONSTANT) return kel;
  const double base = c.adv_total * rpow;
  const double kmx = base / (P.ratio_min * (P.gamma - 1.0) / (P.gamma_p - 1.0) +
                             (P.gamma - 1.0) / (P.gamma_e - 1.0));
  const double kmn = base / (P.ratio_max * (P.gamma - 1.0) / (P.gamma_p - 1.0) +
                             (P.gamma - 1.0) / (P.gamma_e - 1.0));
  return kmax(kmn, kmin(kel, kmx));
}

// heating_model.h:172-179.
__device__ __forceinline__ double clamp_by_ratio(double rmin, double rmax, double total,
                                                 double electron) {
  const double lo = total / (1.0 + rmax);
  const double hi = total / (1.0 + rmin);
  const double b = kmax(0.0, kmin(electron, total));
  return kmax(lo, kmin(b, hi));
}

}  // namespace pangu
