// Device conserved-to-primitive recovery (gfx950, FP64).
//
// Inputs are the "HARM-normalised" 9-vectors of invertor.cc:149-174:
// U = (alpha/sqrt|g|) * conserved with the energy slot holding (tau - D);
// prim carries the previous primitives (initial guess; the 1D schemes take
// the entropy K from it).  Operand order follows the reference schemes.
#pragma once
#include "physics_dev.h"

namespace pangu {

constexpr int MAX_NEWT_ITER = 30;          // recovery/constants.h:15-23
constexpr double NEWT_TOL = 1e-10;
constexpr double MIN_NEWT_TOL = 1e-10;
constexpr int EXTRA_NEWT_ITER = 2;
constexpr double W_TOO_BIG = 1.e20;
constexpr double UTSQ_TOO_BIG = 1.e20;
constexpr double FAIL_VAL = 1.e30;

// Metric view for recovery: full gcov and gcon plus sqrt(-1/gcon00).
struct CenMetric {
  double gcov[16], gcon[16], lapse;
};

struct Noble {
  double Bcon[4], Bcov[4], Qcov[4], Qcon[4], ncon[4];
  double ncov0, Bsq, QdotB, QdotBsq, Qdotn, Qtsq, D;
};

__device__ __forceinline__ void noble_setup(const double U[9], const CenMetric &m, Noble &v) {
  v.Bcon[0] = 0.;
  v.Bcon[1] = U[BX1];
  v.Bcon[2] = U[BX2];
  v.Bcon[3] = U[BX3];
  v.Qcov[0] = U[ENY];
  v.Qcov[1] = U[UX1];
  v.Qcov[2] = U[UX2];
  v.Qcov[3] = U[UX3];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    v.Bcov[r] = m.gcov[4 * r] * v.Bcon[0] + m.gcov[4 * r + 1] * v.Bcon[1] +
                m.gcov[4 * r + 2] * v.Bcon[2] + m.gcov[4 * r + 3] * v.Bcon[3];
    v.Qcon[r] = m.gcon[4 * r] * v.Qcov[0] + m.gcon[4 * r + 1] * v.Qcov[1] +
                m.gcon[4 * r + 2] * v.Qcov[2] + m.gcon[4 * r + 3] * v.Qcov[3];
  }
  v.Bsq = v.Bcon[1] * v.Bcov[1] + v.Bcon[2] * v.Bcov[2] + v.Bcon[3] * v.Bcov[3];
  const double qb = v.Qcov[0] * v.Bcon[0] + v.Qcov[1] * v.Bcon[1] + v.Qcov[2] * v.Bcon[2] +
                    v.Qcov[3] * v.Bcon[3];
  v.QdotB = qb;
  v.QdotBsq = qb * qb;
  // n_mu = (-lapse, 0, 0, 0): n^mu = gcon[mu][0] * (-lapse)
  v.ncov0 = -m.lapse;
#pragma unroll
  for (int r = 0; r < 4; ++r) v.ncon[r] = m.gcon[4 * r] * v.ncov0;
  v.Qdotn = v.Qcon[0] * v.ncov0;
  const double qsq = v.Qcov[0] * v.Qcon[0] + v.Qcov[1] * v.Qcon[1] + v.Qcov[2] * v.Qcon[2] +
                     v.Qcov[3] * v.Qcon[3];
  v.Qtsq = qsq + v.Qdotn * v.Qdotn;
  v.D = U[RHO];
}

__device__ __forceinline__ bool noble_utsq(const double p[9], const CenMetric &m, double &out) {
  double u = 0.;
#pragma unroll
  for (int i = 1; i < 4; i++)
#pragma unroll
    for (int j = 1; j < 4; j++) u += m.gcov[4 * i + j] * p[UX1 + i - 1] * p[UX1 + j - 1];
  if ((u < 0.) && (fabs(u) < 1.0e-13)) u = fabs(u);
  if (u < 0. || u > UTSQ_TOO_BIG) return false;
  out = u;
  return true;
}

__device__ __forceinline__ void noble_velocity(const Noble &v, double W, double Gamma, double p[9]) {
#pragma unroll
  for (int i = 1; i < 4; i++) {
    const double qt = v.Qcon[i] + v.ncon[i] * v.Qdotn;
    p[UX1 + i - 1] = Gamma / (W + v.Bsq) * (qt + v.QdotB * v.Bcon[i] / W);
  }
}

__device__ __forceinline__ double bump_w(double W, double Bsq, double QdotBsq, double Qtsq) {
  int inc = 0;
  while (((W * W * W * (W + 2. * Bsq) - QdotBsq * (2. * W + Bsq)) <= W * W * (Qtsq - Bsq * Bsq)) &&
         (inc < 10)) {
    W *= 10.;
    inc++;
  }
  return W;
}

__device__ __forceinline__ double vsq_of_w(double Bsq, double QdotBsq, double Qtsq, double W) {
  const double Wsq = W * W;
  const double Xsq = (Bsq + W) * (Bsq + W);
  return ((Wsq * Qtsq + QdotBsq * (Bsq + 2. * W)) / (Wsq * Xsq));
}

// ------------------------------------------------- 2D (W, v^2) Newton scheme
// scheme_2d.h:99-404.
__device__ int c2p_2d(const double U[9], double p[9], double gamma, const CenMetric &m, int &iters) {
  iters = 0;
  Noble v;
  noble_setup(U, m, v);
  double utsq;
  if (!noble_utsq(p, m, utsq)) return 2;
  const double gammasq = 1. + utsq;
  double Gamma = sqrt(gammasq);
  double rho0 = v.D / Gamma;
  double uu = p[ENY];
  double pr = ((gamma - 1.) * uu);
  double w = rho0 + uu + pr;
  double W_last = bump_w(w * gammasq, v.Bsq, v.QdotBsq, v.Qtsq);
  double x0 = fabs(W_last);
  double x1;
  {
    const double vs = fabs(vsq_of_w(v.Bsq, v.QdotBsq, v.Qtsq, W_last));
    x1 = (vs > 1.) ? (1.0 - 1.e-15) : vs;
  }
  const double Bsq = v.Bsq, QdotBsq = v.QdotBsq, Qtsq = v.Qtsq, Qdotn = v.Qdotn, D = v.D;
  double f = 1., df = 1., errx = 1.;
  int i_extra = 0, doing_extra = 0, n_iter = 0;
  bool keep = true;
  while (keep) {
    const double W = x0, vsq = x1, Wsq = W * W;
    const double gtmp = 1. - vsq;
    const double p_tmp = ((gamma - 1.) * (W * gtmp - D * sqrt(gtmp)) / gamma);
    const double dPdW = ((gamma - 1.) * (1. - vsq) / gamma);
    const double dPdvsq = ((gamma - 1.) * (0.5 * D / sqrt(1. - vsq) - W) / gamma);
    const double t2 = -0.5 * Bsq + dPdvsq;
    const double t3 = Bsq + W;
    const double t4 = t3 * t3;
    const double t9 = 1.0 / Wsq;
    const double t11 = Qtsq - vsq * t4 + QdotBsq * (Bsq + 2.0 * W) * t9;
    const double t16 = QdotBsq * t9;
    const double t18 = -Qdotn - 0.5 * Bsq * (1.0 + vsq) + 0.5 * t16 - W + p_tmp;
    const double t21 = 1.0 / t3;
    const double t23 = 1.0 / W;
    const double t24 = t16 * t23;
    const double t25 = -1.0 + dPdW - t24;
    const double t35 = t25 * t3 + (Bsq - 2.0 * dPdvsq) * (QdotBsq + vsq * Wsq * W) * t9 * t23;
    const double t36 = 1.0 / t35;
    const double dx0 = -(t2 * t11 + t4 * t18) * t21 * t36;
    const double t40 = (vsq + t24) * t3;
    const double dx1 = -(-t25 * t11 - 2.0 * t40 * t18) * t21 * t36;
    df = -t11 * t11 - t18 * t18;
    f = -0.5 * df;
    const double o0 = x0;
    x0 += dx0;
    x1 += dx1;
    errx = (x0 == 0.) ? fabs(dx0) : fabs(dx0 / x0);
    errx += (x1 == 0.) ? fabs(dx1) : fabs(dx1 / x1);
    x0 = fabs(x0);
    x0 = (x0 > W_TOO_BIG) ? o0 : x0;
    x1 = (x1 < 0.) ? 0. : x1;
    x1 = (x1 > 1.) ? (1. - 1.e-15) : x1;
    if ((fabs(errx) <= NEWT_TOL) && (doing_extra == 0) && (EXTRA_NEWT_ITER > 0)) doing_extra = 1;
    if (doing_extra == 1) i_extra++;
    if (((fabs(errx) <= NEWT_TOL) && (doing_extra == 0)) || (i_extra > EXTRA_NEWT_ITER) ||
        (n_iter >= (MAX_NEWT_ITER - 1)))
      keep = false;
    n_iter++;
  }
  iters = n_iter;
  int retval = 0;
  if (!isfinite(f) || !isfinite(df))
    retval = 2;
  else if (fabs(errx) > MIN_NEWT_TOL)
    retval = 1;
  const double W = x0, vsq = x1;
  if ((retval != 0) || (W == FAIL_VAL)) return retval * 100 + 1;
  if (W <= 0. || W > W_TOO_BIG) return 3;
  if (vsq >= 1.) return 4;
  const double gt = sqrt(1. - vsq);
  Gamma = 1. / gt;
  rho0 = D * gt;
  w = W * (1. - vsq);
  pr = ((gamma - 1.) * (w - rho0) / gamma);
  uu = w - (rho0 + pr);
  if ((rho0 <= 0.) || (uu <= 0.)) return 5;
  p[RHO] = rho0;
  p[ENY] = uu;
  noble_velocity(v, W, Gamma, p);
  return 0;
}

// -------------------------------------------------- 1D v^2 isentropic scheme
// Residual of the 1Dvsq equation (func_1d_gnr, scheme_1d_vsq.h:175-197) at
// v^2 = x for conserved state U, metric m and entropy K, and its scale
// Q~^2 + (B^2 + W)^2 v^2 for relative checks.
__device__ inline double vsq_residual(const double U[9], const CenMetric &m, double K, double gamma,
                                      double x, double *scale);

__device__ __forceinline__ double w_of_vsq(double D, double K, double vsq, double gamma,
                                           double &pr, double &rho, double &uu) {
  const double gtmp = (1. - vsq);
  rho = D * sqrt(gtmp);
  pr = (K * pow(rho, gamma));
  uu = (pr / (gamma - 1.));
  return ((rho + uu + pr) / gtmp);
}

__device__ inline double vsq_residual(const double U[9], const CenMetric &m, double K, double gamma,
                                      double x, double *scale) {
  Noble v;
  noble_setup(U, m, v);
  double pr, rho, uu;
  const double W = w_of_vsq(v.D, K, x, gamma, pr, rho, uu);
  const double fact = v.Bsq + W;
  if (scale) *scale = v.Qtsq + x * fact * fact;
  return v.Qtsq - x * fact * fact + v.QdotBsq * (v.Bsq + 2. * W) / (W * W);
}

// scheme_1d_vsq.h:52-351 (K = previous primitive entropy).  The Newton
// iteration starts from the v^2 of the previous primitives, as the
// reference does, or from x0 when x0 >= 0 (the learned start).
__device__ int c2p_1dvsq(const double U[9], double p[9], double gamma, const CenMetric &m,
                         int &iters, double x0 = -1.0) {
  iters = 0;
  const double K = p[ENT];
  Noble v;
  noble_setup(U, m, v);
  double x;
  if (x0 >= 0.0) {
    x = x0;
  } else {
    double utsq;
    if (!noble_utsq(p, m, utsq)) return 2;
    const double gammasq = 1. + utsq;
    x = 1. - 1. / gammasq;
  }
  const double Bsq = v.Bsq, QdotBsq = v.QdotBsq, Qtsq = v.Qtsq, D = v.D;
  double f = 1., df = 1., errx = 1., W = 0., pr, rho, uu;
  int i_extra = 0, doing_extra = 0, n_iter = 0;
  bool keep = true;
  while (keep) {
    double Wf = w_of_vsq(D, K, x, gamma, pr, rho, uu);
    {
      const double vsq = x;
      const double Wsq = Wf * Wf;
      const double W3 = Wf * Wsq;
      const double dWdvsq =
          ((gamma * (2. - gamma) * pr + (gamma - 1.) * rho) / (2. * (gamma - 1.) * (1. - vsq) * (1. - vsq)));
      const double fact = (Bsq + Wf);
      const double resid = Qtsq - vsq * fact * fact + QdotBsq * (Bsq + 2. * Wf) / Wsq;
      const double jac = -fact * (fact + 2. * dWdvsq * (vsq + QdotBsq / W3));
      const double dx = -resid / jac;
      f = 0.5 * resid * resid;
      df = -2. * f;
      const double xo = x;
      x += dx;
      x = (x >= 1.0) ? (0.5 * (xo + 1.)) : x;
      x = (x < -1.e-10) ? (0.5 * xo) : x;
      x = fabs(x);
      const double Wo = W;
      W = w_of_vsq(D, K, x, gamma, pr, rho, uu);
      errx = (W == 0.) ? fabs(W - Wo) : fabs((W - Wo) / W);
      errx += (x == 0.) ? fabs(x - xo) : fabs((x - xo) / x);
    }
    if ((fabs(errx) <= NEWT_TOL) && (doing_extra == 0) && (EXTRA_NEWT_ITER > 0)) doing_extra = 1;
    if (doing_extra == 1) i_extra++;
    if (((fabs(errx) <= NEWT_TOL) && (doing_extra == 0)) || (i_extra > EXTRA_NEWT_ITER) ||
        (n_iter >= (MAX_NEWT_ITER - 1)))
      keep = false;
    n_iter++;
  }
  iters = n_iter;
  int retval = 0;
  if (!isfinite(f) || !isfinite(df))
    retval = 2;
  else if (fabs(errx) > MIN_NEWT_TOL)
    retval = 1;
  if (retval != 0) return retval * 100 + 1;
  const double vsq = x;
  if ((vsq >= 1.) || (vsq < 0.)) return 4;
  double rho0;
  W = w_of_vsq(D, K, vsq, gamma, pr, rho0, uu);
  const double Gamma = 1. / sqrt(1. - vsq);
  if ((rho0 <= 0.) || (uu <= 0.)) return 5;
  p[RHO] = rho0;
  p[ENY] = uu;
  noble_velocity(v, W, Gamma, p);
  return 0;
}

// --------------------------------------------------- 1D W isentropic scheme
// scheme_1d.h:196-568 with func_1d_orig2.
__device__ int c2p_1d(const double U[9], double p[9], double gamma, const CenMetric &m, int &iters) {
  iters = 0;
  const double K = p[ENT];
  Noble v;
  noble_setup(U, m, v);
  double utsq;
  if (!noble_utsq(p, m, utsq)) return 2;
  const double uvsq = 1. + utsq;
  double Gamma = sqrt(uvsq);
  double rho0 = v.D / Gamma;
  double pr = (K * pow(rho0, gamma));
  double uu = (pr / (gamma - 1.));
  const double w = rho0 + uu + pr;
  double x = bump_w(w * uvsq, v.Bsq, v.QdotBsq, v.Qtsq);
  const double Bsq = v.Bsq, QdotBsq = v.QdotBsq, Qtsq = v.Qtsq, D = v.D;
  double f = 1., df = 1., errx = 1.;
  int i_extra = 0, doing_extra = 0, n_iter = 0;
  bool keep = true;
  while (keep) {
    const double W = x;
    const double t1 = D * D;
    const double t2 = QdotBsq * t1;
    const double t3 = t2 * Bsq;
    const double t5 = Bsq * Bsq;
    const double t8 = t1 * Bsq;
    const double t10 = t1 * W;
    const double t21 = W * W;
    const double rho = t1 * (1. + gamma * K / (gamma - 1.)) / W;
    const double t23 = rho * rho;
    const double t26 = 1 / t1;
    const double resid = (t3 + (2.0 * t2 + ((Qtsq - t5) * t1 + (-2.0 * t8 - t10) * W) * W) * W +
                          (t5 + (2.0 * Bsq + W) * W) * t21 * t23) *
                         t26 / t21;
    const double jac =
        -2.0 * (t3 + (t2 + (t8 + t10) * t21) * W + (t5 + Bsq * W) * t21 * t23) * t26 / t21 / W;
    const double dx = -resid / jac;
    f = 0.5 * resid * resid;
    df = -2. * f;
    x += dx;
    int inc = 0;
    while (((x * x * x * (x + 2. * Bsq) - QdotBsq * (2. * x + Bsq)) <= x * x * (Qtsq - Bsq * Bsq)) &&
           (inc < 10)) {
      x -= (1. * inc) * dx / 10.;
      inc++;
    }
    errx = (x == 0.) ? fabs(dx) : fabs(dx / x);
    x = fabs(x);
    if ((fabs(errx) <= NEWT_TOL) && (doing_extra == 0) && (EXTRA_NEWT_ITER > 0)) doing_extra = 1;
    if (doing_extra == 1) i_extra++;
    if (((fabs(errx) <= NEWT_TOL) && (doing_extra == 0)) || (i_extra > EXTRA_NEWT_ITER) ||
        (n_iter >= (MAX_NEWT_ITER - 1)))
      keep = false;
    n_iter++;
  }
  iters = n_iter;
  int retval = 0;
  if (!isfinite(f) || !isfinite(df) || !isfinite(x))
    retval = 2;
  else if (fabs(errx) > MIN_NEWT_TOL)
    retval = 1;
  const double W = x;
  if ((retval != 0) || (W == FAIL_VAL)) return retval * 100 + 1;
  if (W <= 0. || W > W_TOO_BIG) return 3;
  const double vsq = vsq_of_w(Bsq, QdotBsq, Qtsq, W);
  if (vsq >= 1.) return 4;
  const double gt = sqrt(1. - vsq);
  Gamma = 1. / gt;
  rho0 = D * gt;
  pr = (K * pow(rho0, gamma));
  uu = (pr / (gamma - 1.));
  if ((rho0 <= 0.) || (uu <= 0.)) return 5;
  p[RHO] = rho0;
  p[ENY] = uu;
  noble_velocity(v, W, Gamma, p);
  return 0;
}

// ----------------------------------------------------- Kastaun et al. 2021
// scheme_kastaun.h:84-108.
__device__ __forceinline__ double kastaun_root(double mu, double D, double q, double bsq, double rsq,
                                               double rbsq, double gamma) {
  const double x = 1.0 / (1.0 + mu * bsq);
  const double xsq = x * x;
  const double rbarsq = x * (rsq * x + mu * (x + 1.0) * rbsq);
  const double musq = mu * mu;
  double vsq = musq * rbarsq;
  vsq = kmin(kmax(vsq, 0.0), 1.0 - 1.0e-14);
  const double il = sqrt(1.0 - vsq);
  const double qbar = q - 0.5 * bsq - 0.5 * musq * xsq * (bsq * rsq - rbsq);
  const double eod = qbar - mu * rbarsq + 1.0;
  const double rho = D * il;
  double eps = D * eod - rho;
  eps = kmax(eps, 1.0e-300);
  const double pr = (gamma - 1.0) * eps;
  const double hh = 1.0 + (eps + pr) / rho;
  const double nu = kmax(hh * il, eod + pr / D);
  return mu - 1.0 / (nu + mu * rbarsq);
}

// Kastaun inputs (scheme_kastaun.h:167-213): D, q = tau/D, r^2, b^2, (r.b)^2
// and what the primitive reconstruction needs.
struct KastaunIn {
  double D, q, bsq, rsq, rbsq, rb, inv_D, sqrt_D;
  double Su[3], B[3];
};

// Returns 0, or 1..4 for an invalid metric, D, tau or S/B invariants.
__device__ __forceinline__ int kastaun_setup(const double U[9], const CenMetric &m, KastaunIn &k) {
  const double g0 = m.gcov[5], g1 = m.gcov[6], g2 = m.gcov[7], g3 = m.gcov[10], g4 = m.gcov[11],
               g5 = m.gcov[15];
  const double det = g0 * g3 * g5 + 2.0 * g1 * g2 * g4 - g0 * g4 * g4 - g1 * g1 * g5 - g2 * g2 * g3;
  if (!(det > 0.0) || !isfinite(det)) return 1;
  const double idet = 1.0 / det;
  const double i0 = (g3 * g5 - g4 * g4) * idet, i1 = (g2 * g4 - g1 * g5) * idet,
               i2 = (g1 * g4 - g2 * g3) * idet, i3 = (g0 * g5 - g2 * g2) * idet,
               i4 = (g1 * g2 - g0 * g4) * idet, i5 = (g0 * g3 - g1 * g1) * idet;
  const double D = U[RHO];
  if (!(D > 0.0) || !isfinite(D)) return 2;
  const double ncov0 = -m.lapse;
  const double Qcon0 =
      m.gcon[0] * U[ENY] + m.gcon[1] * U[UX1] + m.gcon[2] * U[UX2] + m.gcon[3] * U[UX3];
  const double Qdotn = Qcon0 * ncov0;
  const double tau = -Qdotn - D;
  if (!isfinite(tau)) return 3;
  const double S0 = U[UX1], S1 = U[UX2], S2 = U[UX3];
  k.Su[0] = i0 * S0 + i1 * S1 + i2 * S2;
  k.Su[1] = i1 * S0 + i3 * S1 + i4 * S2;
  k.Su[2] = i2 * S0 + i4 * S1 + i5 * S2;
  const double B0 = U[BX1], B1 = U[BX2], B2 = U[BX3];
  k.B[0] = B0;
  k.B[1] = B1;
  k.B[2] = B2;
  const double Bd0 = g0 * B0 + g1 * B1 + g2 * B2;
  const double Bd1 = g1 * B0 + g3 * B1 + g4 * B2;
  const double Bd2 = g2 * B0 + g4 * B1 + g5 * B2;
  const double Bsq = B0 * Bd0 + B1 * Bd1 + B2 * Bd2;
  const double Ssq = k.Su[0] * S0 + k.Su[1] * S1 + k.Su[2] * S2;
  const double SdotB = B0 * S0 + B1 * S1 + B2 * S2;
  if (!isfinite(Bsq) || !isfinite(Ssq) || !isfinite(SdotB)) return 4;
  k.D = D;
  k.inv_D = 1.0 / D;
  k.sqrt_D = sqrt(D);
  k.bsq = Bsq * k.inv_D;
  k.rsq = Ssq * k.inv_D * k.inv_D;
  k.rb = SdotB / (D * k.sqrt_D);
  k.rbsq = k.rb * k.rb;
  k.q = tau * k.inv_D;
  return 0;
}

// Illinois false position on [lo, hi] with end values flo, fhi
// (scheme_kastaun.h:110-165).  Returns 0 with the root in mu, 1 non-finite
// end values, 2 no sign change, 3 non-finite iterate, 4 no convergence in
// MAX_NEWT_ITER.  iters counts the false-position steps.
__device__ __forceinline__ int kastaun_iterate(const KastaunIn &k, double gamma, double lo, double flo,
                                               double hi, double fhi, double &mu, int &iters) {
  const double D = k.D, q = k.q, bsq = k.bsq, rsq = k.rsq, rbsq = k.rbsq;
  iters = 0;
  mu = 0.0;
  if (!isfinite(flo) || !isfinite(fhi)) return 1;
  if (fabs(flo) <= NEWT_TOL) {
    mu = lo;
    return 0;
  }
  if (fabs(fhi) <= NEWT_TOL) {
    mu = hi;
    return 0;
  }
  if (flo * fhi > 0.0) return 2;
  int side = 0;
  mu = lo;
  for (int it = 0; it < MAX_NEWT_ITER; ++it) {
    const double old = mu;
    mu = (fhi * lo - flo * hi) / (fhi - flo);
    const double fm = kastaun_root(mu, D, q, bsq, rsq, rbsq, gamma);
    iters = it + 1;
    if (!isfinite(mu) || !isfinite(fm)) return 3;
    if (fabs((mu - old) / (mu + 1.0e-300)) <= NEWT_TOL) return 0;
    if (fm * flo >= 0.0) {
      if (side == 1) {
        const double wgt = 1.0 - fm / flo;
        fhi = (wgt > 0.0) ? fhi * wgt : 0.5 * fhi;
      }
      lo = mu;
      flo = fm;
      side = 1;
    } else {
      if (side == -1) {
        const double wgt = 1.0 - fm / fhi;
        flo = (wgt > 0.0) ? flo * wgt : 0.5 * flo;
      }
      hi = mu;
      fhi = fm;
      side = -1;
    }
  }
  return 4;
}

__device__ __forceinline__ int kastaun_solve(const KastaunIn &k, double gamma, double lo, double hi,
                                             double &mu, int &iters) {
  const double flo = kastaun_root(lo, k.D, k.q, k.bsq, k.rsq, k.rbsq, gamma);
  const double fhi = kastaun_root(hi, k.D, k.q, k.bsq, k.rsq, k.rbsq, gamma);
  return kastaun_iterate(k, gamma, lo, flo, hi, fhi, mu, iters);
}

// Primitives from the root (scheme_kastaun.h:215-259); 0, 5 (v >= 1) or
// 6 (rho or eps not positive).
__device__ __forceinline__ int kastaun_finish(const KastaunIn &k, double mu, double p[9]) {
  const double bsq = k.bsq, rsq = k.rsq, rbsq = k.rbsq, q = k.q, D = k.D;
  const double x = 1.0 / (1.0 + mu * bsq);
  const double rbarsq = x * (rsq * x + mu * (x + 1.0) * rbsq);
  const double vsq = mu * mu * rbarsq;
  if (!(vsq >= 0.0) || vsq >= 1.0 || !isfinite(vsq)) return 5;
  const double lorentz = 1.0 / sqrt(1.0 - vsq);
  const double rho = D / lorentz;
  const double qbar = q - 0.5 * bsq - 0.5 * mu * mu * x * x * (bsq * rsq - rbsq);
  const double eod = qbar - mu * rbarsq + 1.0;
  const double eps = D * eod - rho;
  if (!(rho > 0.0) || !(eps > 0.0) || !isfinite(eps)) return 6;
  const double Wmux = lorentz * mu * x;
  p[RHO] = rho;
  p[ENY] = eps;
  p[UX1] = Wmux * (k.Su[0] * k.inv_D + k.rb * mu * (k.B[0] / k.sqrt_D));
  p[UX2] = Wmux * (k.Su[1] * k.inv_D + k.rb * mu * (k.B[1] / k.sqrt_D));
  p[UX3] = Wmux * (k.Su[2] * k.inv_D + k.rb * mu * (k.B[2] / k.sqrt_D));
  return 0;
}

// The search and reconstruction of c2p_kastaun from a completed setup.
__device__ __forceinline__ int kastaun_plain_k(const KastaunIn &k, double gamma, double p[9],
                                               int &iters, double *mu_out = nullptr) {
  double mu;
  const int rs = kastaun_solve(k, gamma, 0.0, 1.0, mu, iters);
  if (rs != 0) return 100 + rs;
  if (mu_out) *mu_out = mu;
  return kastaun_finish(k, mu, p);
}

// The five dimensionless invariants the learned bracket's network reads:
// log q, log r^2, log b^2, log (r.b)^2 (each floored: 1e-12, 1e-24, 1e-24,
// 1e-48) and (r.b)^2 / (r^2 b^2) clamped to [0, 1].  Logarithms resolve the
// cold, slow and weakly magnetised states, where 1 - mu scales with q and
// the plain values sit within rounding of 0.
__device__ inline void c2p_features(const KastaunIn &k, float f[5]) {
  f[0] = (float)log(k.q > 1e-12 ? k.q : 1e-12);
  f[1] = (float)log(k.rsq > 1e-24 ? k.rsq : 1e-24);
  f[2] = (float)log(k.bsq > 1e-24 ? k.bsq : 1e-24);
  f[3] = (float)log(k.rbsq > 1e-48 ? k.rbsq : 1e-48);
  double c = k.rbsq / (k.rsq * k.bsq + 1e-300);
  c = c < 0.0 ? 0.0 : (c > 1.0 ? 1.0 : c);
  f[4] = (float)c;
}

// scheme_kastaun.h:110-259: Illinois false position on [0,1].  Status 0,
// 1..4 from the setup, 101..104 from the root search, 5..6 from the
// reconstruction.  mu_out (optional) receives the root.
__device__ int c2p_kastaun(const double U[9], double p[9], double gamma, const CenMetric &m,
                           int &iters, double *mu_out = nullptr) {
  iters = 0;
  KastaunIn k;
  if (int e = kastaun_setup(U, m, k)) return e;
  double mu;
  const int rs = kastaun_solve(k, gamma, 0.0, 1.0, mu, iters);
  if (rs != 0) return 100 + rs;
  if (mu_out) *mu_out = mu;
  return kastaun_finish(k, mu, p);
}

// Kastaun with a learned starting bracket.  The master function rises
// through its root, so the sign of f(g) at the guess g says on which side the
// root lies: the bracket is [g, g (1 + delta)] (or [g (1 - delta), g]), widened
// by 4x up to three times while it holds no sign change, and [0, 1] when that
// fails too.  The root found is a root of the same function; only the
// false-position path differs from c2p_kastaun (agreement to the 1e-10 step
// tolerance, not bit for bit).  iters counts false-position steps, evals
// every evaluation of the master function.
__device__ __forceinline__ int kastaun_guess_k(const KastaunIn &k, double gamma, double p[9],
                                               double guess, double delta, int &iters, int &evals,
                                               double *mu_out = nullptr) {
  iters = 0;
  evals = 0;
  double mu = 0.0;
  int rs = 2;
  if (guess > 0.0 && guess < 1.0) {
    double x0 = guess;
    double f0 = kastaun_root(x0, k.D, k.q, k.bsq, k.rsq, k.rbsq, gamma);
    evals = 1;
    if (isfinite(f0) && fabs(f0) <= NEWT_TOL) {
      mu = x0;
      rs = 0;
    } else if (isfinite(f0)) {
      const double dir = (f0 < 0.0) ? 1.0 : -1.0;
      double dl = delta;
      for (int e = 0; e < 4; ++e) {
        const double x1 = kmin(1.0, kmax(0.0, x0 * (1.0 + dir * dl)));
        const double f1 = kastaun_root(x1, k.D, k.q, k.bsq, k.rsq, k.rbsq, gamma);
        evals += 1;
        if (!isfinite(f1)) break;
        if (f0 * f1 <= 0.0) {
          int it = 0;
          if (x0 < x1)
            rs = kastaun_iterate(k, gamma, x0, f0, x1, f1, mu, it);
          else
            rs = kastaun_iterate(k, gamma, x1, f1, x0, f0, mu, it);
          iters = it;
          evals += it;
          break;
        }
        if (x1 == 0.0 || x1 == 1.0) break;
        x0 = x1;
        f0 = f1;
        dl *= 4.0;
      }
    }
  }
  if (rs != 0) {
    int it2 = 0;
    rs = kastaun_solve(k, gamma, 0.0, 1.0, mu, it2);
    iters += it2;
    evals += 2 + it2;
  }
  if (rs != 0) return 100 + rs;
  if (mu_out) *mu_out = mu;
  return kastaun_finish(k, mu, p);
}

__device__ int c2p_kastaun_guess(const double U[9], double p[9], double gamma, const CenMetric &m,
                                 double guess, double delta, int &iters, int &evals,
                                 double *mu_out = nullptr) {
  iters = 0;
  evals = 0;
  KastaunIn k;
  if (int e = kastaun_setup(U, m, k)) return e;
  return kastaun_guess_k(k, gamma, p, guess, delta, iters, evals, mu_out);
}

}  // namespace pangu
