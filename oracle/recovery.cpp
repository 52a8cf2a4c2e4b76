// Oracle: conserved-to-primitive recovery schemes.  TEST INFRASTRUCTURE ONLY.
//
// U and prim are the "HARM-normalised" 9-vectors built by the invertor
// (invertor.cc:149-174): U = (alpha/sqrt|g|) * conserved with the energy slot
// holding (tau - D), prim holds the previous primitives as the initial guess.
#include "oracle.h"

namespace oracle {

static void lower4(const real v[4], const real g[4][4], real out[4]) {
  for (int i = 0; i < 4; ++i) {
    real acc = 0;
    for (int j = 0; j < 4; ++j) acc += g[i][j] * v[j];
    out[i] = acc;
  }
}

// Shared set-up of the Noble et al. schemes (scheme_2d.h:271-308,
// scheme_1d_vsq.h:234-271, scheme_1d.h:423-460).
struct NobleInv {
  real Bcon[4], Bcov[4], Qcov[4], Qcon[4], ncov[4], ncon[4];
  real Bsq, QdotB, QdotBsq, Qdotn, Qsq, Qtsq, D;
};
static void noble_setup(const real U[NPRIM_RECV], const real gcov[4][4], const real gcon[4][4],
                        NobleInv &v) {
  v.Bcon[0] = 0.;
  v.Bcon[1] = U[BX1];
  v.Bcon[2] = U[BX2];
  v.Bcon[3] = U[BX3];
  lower4(v.Bcon, gcov, v.Bcov);
  v.Qcov[0] = U[ENY];
  v.Qcov[1] = U[UX1];
  v.Qcov[2] = U[UX2];
  v.Qcov[3] = U[UX3];
  lower4(v.Qcov, gcon, v.Qcon);  // raise_g: same contraction with gcon
  real bsq = 0.;
  for (int i = 1; i < 4; ++i) bsq += v.Bcon[i] * v.Bcov[i];
  v.Bsq = bsq;
  real qb = 0.;
  for (int i = 0; i < 4; ++i) qb += v.Qcov[i] * v.Bcon[i];
  v.QdotB = qb;
  v.QdotBsq = qb * qb;
  const real lapse = std::sqrt(-1.0 / gcon[0][0]);
  for (int i = 0; i < 4; ++i) v.ncov[i] = -lapse * (real)(i == 0);
  lower4(v.ncov, gcon, v.ncon);
  v.Qdotn = v.Qcon[0] * v.ncov[0];
  real qsq = 0.;
  for (int i = 0; i < 4; ++i) qsq += v.Qcov[i] * v.Qcon[i];
  v.Qsq = qsq;
  v.Qtsq = qsq + v.Qdotn * v.Qdotn;
  v.D = U[RHO];
}

// gamma_ij u~^i u~^j over the old primitives; returns false for the
// reference's retval 2 (scheme_2d.h:311-323).
static bool noble_utsq(const real prim[NPRIM_RECV], const real gcov[4][4], real *utsq_out) {
  real utsq = 0.;
  for (int i = 1; i < 4; i++)
    for (int j = 1; j < 4; j++) utsq += gcov[i][j] * prim[UX1 + i - 1] * prim[UX1 + j - 1];
  if ((utsq < 0.) && (std::fabs(utsq) < 1.0e-13)) utsq = std::fabs(utsq);
  if (utsq < 0. || utsq > UTSQ_TOO_BIG) return false;
  *utsq_out = utsq;
  return true;
}

static void noble_velocity(const NobleInv &v, real W, real Gamma, real prim[NPRIM_RECV]) {
  real Qtcon[4];
  for (int i = 1; i < 4; i++) Qtcon[i] = v.Qcon[i] + v.ncon[i] * v.Qdotn;
  for (int i = 1; i < 4; i++)
    prim[UX1 + i - 1] = Gamma / (W + v.Bsq) * (Qtcon[i] + v.QdotB * v.Bcon[i] / W);
}

// Bump W by 10x until it lies above the v^2 < 1 bound (scheme_2d.h:339-346).
static real bump_w(real W, real Bsq, real QdotBsq, real Qtsq) {
  int inc = 0;
  while (((W * W * W * (W + 2. * Bsq) - QdotBsq * (2. * W + Bsq)) <= W * W * (Qtsq - Bsq * Bsq)) &&
         (inc < 10)) {
    W *= 10.;
    inc++;
  }
  return W;
}

// ---------------------------------------------------------------- 2D scheme
static real vsq_of_w(real Bsq, real QdotBsq, real Qtsq, real W) {
  const real Wsq = W * W;
  const real Xsq = (Bsq + W) * (Bsq + W);
  return ((Wsq * Qtsq + QdotBsq * (Bsq + 2. * W)) / (Wsq * Xsq));
}

// Newton step of the (W, v^2) system (scheme_2d.h:99-156).
static void func_vsq(real Bsq, real QdotBsq, real Qtsq, real Qdotn, real D, const real x[2],
                     real dx[2], real *f, real *df, real gamma) {
  const real W = x[0];
  const real vsq = x[1];
  const real Wsq = W * W;
  const real gtmp = 1. - vsq;
  const real p_tmp = ((gamma - 1.) * (W * gtmp - D * std::sqrt(gtmp)) / gamma);
  const real dPdW = ((gamma - 1.) * (1. - vsq) / gamma);
  const real dPdvsq = ((gamma - 1.) * (0.5 * D / std::sqrt(1. - vsq) - W) / gamma);
  const real t2 = -0.5 * Bsq + dPdvsq;
  const real t3 = Bsq + W;
  const real t4 = t3 * t3;
  const real t9 = 1.0 / Wsq;
  const real t11 = Qtsq - vsq * t4 + QdotBsq * (Bsq + 2.0 * W) * t9;
  const real t16 = QdotBsq * t9;
  const real t18 = -Qdotn - 0.5 * Bsq * (1.0 + vsq) + 0.5 * t16 - W + p_tmp;
  const real t21 = 1.0 / t3;
  const real t23 = 1.0 / W;
  const real t24 = t16 * t23;
  const real t25 = -1.0 + dPdW - t24;
  const real t35 = t25 * t3 + (Bsq - 2.0 * dPdvsq) * (QdotBsq + vsq * Wsq * W) * t9 * t23;
  const real t36 = 1.0 / t35;
  dx[0] = -(t2 * t11 + t4 * t18) * t21 * t36;
  const real t40 = (vsq + t24) * t3;
  dx[1] = -(-t25 * t11 - 2.0 * t40 * t18) * t21 * t36;
  const real r0 = t11, r1 = t18;
  *df = -r0 * r0 - r1 * r1;
  *f = -0.5 * (*df);
}

// Newton-Raphson driver with two extra iterations after convergence
// (scheme_2d.h:158-255).
static int gnr_2d(real Bsq, real QdotBsq, real Qtsq, real Qdotn, real D, real x[2], real gamma,
                  int *iters) {
  real f = 1., df = 1., dx[2], x_old[2];
  real errx = 1.;
  int i_extra = 0, doing_extra = 0, n_iter = 0;
  bool keep = true;
  while (keep) {
    func_vsq(Bsq, QdotBsq, Qtsq, Qdotn, D, x, dx, &f, &df, gamma);
    x_old[0] = x[0];
    x_old[1] = x[1];
    x[0] += dx[0];
    x[1] += dx[1];
    errx = (x[0] == 0.) ? std::fabs(dx[0]) : std::fabs(dx[0] / x[0]);
    errx += (x[1] == 0.) ? std::fabs(dx[1]) : std::fabs(dx[1] / x[1]);
    // validate_x (scheme_2d.h:86-97)
    x[0] = std::fabs(x[0]);
    x[0] = (x[0] > W_TOO_BIG) ? x_old[0] : x[0];
    x[1] = (x[1] < 0.) ? real(0.) : x[1];
    x[1] = (x[1] > 1.) ? real(1. - 1.e-15) : x[1];
    if ((std::fabs(errx) <= NEWT_TOL) && (doing_extra == 0) && (EXTRA_NEWT_ITER > 0)) doing_extra = 1;
    if (doing_extra == 1) i_extra++;
    if (((std::fabs(errx) <= NEWT_TOL) && (doing_extra == 0)) || (i_extra > EXTRA_NEWT_ITER) ||
        (n_iter >= (MAX_NEWT_ITER - 1)))
      keep = false;
    n_iter++;
  }
  *iters = n_iter;
  if (!std::isfinite(f) || !std::isfinite(df)) return 2;
  if (std::fabs(errx) > MIN_NEWT_TOL) return 1;
  return 0;
}

// Scheme2D::invert (scheme_2d.h:257-404).
int c2p_2d(real U[NPRIM_RECV], real prim[NPRIM_RECV], real gamma, const real gcov[4][4],
           const real gcon[4][4], int *iters) {
  *iters = 0;
  prim[BX1] = U[BX1];
  prim[BX2] = U[BX2];
  prim[BX3] = U[BX3];
  NobleInv v;
  noble_setup(U, gcov, gcon, v);
  real utsq;
  if (!noble_utsq(prim, gcov, &utsq)) return 2;
  const real gammasq = 1. + utsq;
  real Gamma = std::sqrt(gammasq);
  real rho0 = v.D / Gamma;
  real uu = prim[ENY];
  real p = ((gamma - 1.) * uu);
  real w = rho0 + uu + p;
  real W_last = w * gammasq;
  W_last = bump_w(W_last, v.Bsq, v.QdotBsq, v.Qtsq);
  real x[2];
  x[0] = std::fabs(W_last);
  {
    const real vs = std::fabs(vsq_of_w(v.Bsq, v.QdotBsq, v.Qtsq, W_last));
    x[1] = (vs > 1.) ? real(1.0 - 1.e-15) : vs;
  }
  int retval = gnr_2d(v.Bsq, v.QdotBsq, v.Qtsq, v.Qdotn, v.D, x, gamma, iters);
  const real W = x[0];
  const real vsq = x[1];
  if ((retval != 0) || (W == FAIL_VAL)) return retval * 100 + 1;
  if (W <= 0. || W > W_TOO_BIG) return 3;
  if (vsq >= 1.) return 4;
  const real gtmp = std::sqrt(1. - vsq);
  Gamma = 1. / gtmp;
  rho0 = v.D * gtmp;
  w = W * (1. - vsq);
  p = ((gamma - 1.) * (w - rho0) / gamma);
  uu = w - (rho0 + p);
  if ((rho0 <= 0.) || (uu <= 0.)) return 5;
  prim[RHO] = rho0;
  prim[ENY] = uu;
  prim[ENT] = (gamma - 1.) * uu * std::pow(rho0, -gamma);
  noble_velocity(v, W, Gamma, prim);
  prim[BX1] = U[BX1];
  prim[BX2] = U[BX2];
  prim[BX3] = U[BX3];
  return 0;
}

// ---------------------------------------------------------- 1D v^2 scheme
static real w_of_vsq(real D, real K, real vsq, real *p, real *rho, real *uu, real gamma) {
  const real gtmp = (1. - vsq);
  *rho = D * std::sqrt(gtmp);
  *p = (K * std::pow(*rho, gamma));
  *uu = (*p / (gamma - 1.));
  return ((*rho + *uu + *p) / gtmp);
}

// scheme_1d_vsq.h:91-117.
static void func_1d_vsq(real Bsq, real QdotBsq, real Qtsq, real D, real K, const real x[1],
                        real dx[1], real *f, real *df, real gamma) {
  real p, rho, uu;
  const real vsq = x[0];
  const real W = w_of_vsq(D, K, vsq, &p, &rho, &uu, gamma);
  const real Wsq = W * W;
  const real W3 = W * Wsq;
  const real dWdvsq = ((gamma * (2. - gamma) * p + (gamma - 1.) * rho) /
                       (2. * (gamma - 1.) * (1. - vsq) * (1. - vsq)));
  const real fact = (Bsq + W);
  const real resid = Qtsq - vsq * fact * fact + QdotBsq * (Bsq + 2. * W) / Wsq;
  const real jac = -fact * (fact + 2. * dWdvsq * (vsq + QdotBsq / W3));
  dx[0] = -resid / jac;
  *f = 0.5 * resid * resid;
  *df = -2. * (*f);
}

// The residual func_1d_vsq computes (scheme_1d_vsq.h:91-117) at v^2 = vsq,
// for the recovery-normalised state U and entropy K, and its scale
// Q~^2 + vsq (B^2 + W)^2 (the two terms it balances).  An independent check
// of a 1Dvsq result.
void vsq_residual(const real U[NPRIM_RECV], real K, real vsq, real gamma, const real gcov[4][4],
                  const real gcon[4][4], real *res, real *scale) {
  NobleInv v;
  noble_setup(U, gcov, gcon, v);
  real p, rho, uu;
  const real W = w_of_vsq(v.D, K, vsq, &p, &rho, &uu, gamma);
  const real fact = v.Bsq + W;
  *res = v.Qtsq - vsq * fact * fact + v.QdotBsq * (v.Bsq + 2. * W) / (W * W);
  *scale = v.Qtsq + vsq * fact * fact;
}

// scheme_1d_vsq.h:119-220.
static int gnr_1dvsq(real Bsq, real QdotBsq, real Qtsq, real D, real K, real x[1], real gamma,
                     int *iters) {
  real f = 1., df = 1., dx[1], x_old[1];
  real errx = 1.;
  real W = 0., W_old = 0., p, rho, uu;
  int i_extra = 0, doing_extra = 0, n_iter = 0;
  bool keep = true;
  while (keep) {
    func_1d_vsq(Bsq, QdotBsq, Qtsq, D, K, x, dx, &f, &df, gamma);
    x_old[0] = x[0];
    x[0] += dx[0];
    // validate_x (scheme_1d_vsq.h:52-61)
    x[0] = (x[0] >= 1.0) ? (0.5 * (x_old[0] + 1.)) : x[0];
    x[0] = (x[0] < -1.e-10) ? (0.5 * x_old[0]) : x[0];
    x[0] = std::fabs(x[0]);
    W_old = W;
    W = w_of_vsq(D, K, x[0], &p, &rho, &uu, gamma);
    errx = (W == 0.) ? std::fabs(W - W_old) : std::fabs((W - W_old) / W);
    errx += (x[0] == 0.) ? std::fabs(x[0] - x_old[0]) : std::fabs((x[0] - x_old[0]) / x[0]);
    if ((std::fabs(errx) <= NEWT_TOL) && (doing_extra == 0) && (EXTRA_NEWT_ITER > 0)) doing_extra = 1;
    if (doing_extra == 1) i_extra++;
    if (((std::fabs(errx) <= NEWT_TOL) && (doing_extra == 0)) || (i_extra > EXTRA_NEWT_ITER) ||
        (n_iter >= (MAX_NEWT_ITER - 1)))
      keep = false;
    n_iter++;
  }
  *iters = n_iter;
  if (!std::isfinite(f) || !std::isfinite(df)) return 2;
  if (std::fabs(errx) > MIN_NEWT_TOL) return 1;
  return 0;
}

// Scheme1Dvsq::invert (scheme_1d_vsq.h:222-351); the entropy K is the
// previous primitive entropy (:230).
int c2p_1dvsq(real U[NPRIM_RECV], real prim[NPRIM_RECV], real gamma, const real gcov[4][4],
              const real gcon[4][4], int *iters) {
  *iters = 0;
  const real K = prim[ENT];
  prim[BX1] = U[BX1];
  prim[BX2] = U[BX2];
  prim[BX3] = U[BX3];
  NobleInv v;
  noble_setup(U, gcov, gcon, v);
  real utsq;
  if (!noble_utsq(prim, gcov, &utsq)) return 2;
  const real gammasq = 1. + utsq;
  real x[1];
  x[0] = 1. - 1. / gammasq;
  const int retval = gnr_1dvsq(v.Bsq, v.QdotBsq, v.Qtsq, v.D, K, x, gamma, iters);
  if (retval != 0) return retval * 100 + 1;
  const real vsq = x[0];
  if ((vsq >= 1.) || (vsq < 0.)) return 4;
  real p, rho0, uu;
  const real W = w_of_vsq(v.D, K, vsq, &p, &rho0, &uu, gamma);
  const real gtmp = std::sqrt(1. - vsq);
  const real Gamma = 1. / gtmp;
  if ((rho0 <= 0.) || (uu <= 0.)) return 5;
  prim[RHO] = rho0;
  prim[ENY] = uu;
  prim[ENT] = (gamma - 1.) * uu * std::pow(rho0, -gamma);
  noble_velocity(v, W, Gamma, prim);
  prim[BX1] = U[BX1];
  prim[BX2] = U[BX2];
  prim[BX3] = U[BX3];
  return 0;
}

// ------------------------------------------------------------- 1D W scheme
// func_1d_orig2 (scheme_1d.h:376-409); USE_ISENTROPIC is undefined.
static void func_1d_orig2(real QdotBsq, real K, real D, real Bsq, real Qtsq, const real x[1],
                          real dx[1], real *f, real *df, real gamma) {
  const real W = x[0];
  const real t1 = D * D;
  const real t2 = QdotBsq * t1;
  const real t3 = t2 * Bsq;
  const real t5 = Bsq * Bsq;
  const real t8 = t1 * Bsq;
  const real t10 = t1 * W;
  const real t21 = W * W;
  const real rho = t1 * (1. + gamma * K / (gamma - 1.)) / W;
  const real t23 = rho * rho;
  const real t26 = 1 / t1;
  const real resid =
      (t3 + (2.0 * t2 + ((Qtsq - t5) * t1 + (-2.0 * t8 - t10) * W) * W) * W +
       (t5 + (2.0 * Bsq + W) * W) * t21 * t23) *
      t26 / t21;
  const real jac = -2.0 * (t3 + (t2 + (t8 + t10) * t21) * W + (t5 + Bsq * W) * t21 * t23) * t26 /
                   t21 / W;
  dx[0] = -resid / jac;
  *f = 0.5 * resid * resid;
  *df = -2. * (*f);
}

// scheme_1d.h:196-304.
static int gnr_1d(real QdotBsq, real K, real D, real Bsq, real Qtsq, real x[1], real gamma,
                  int *iters) {
  real f = 1., df = 1., dx[1];
  real errx = 1.;
  int i_extra = 0, doing_extra = 0, n_iter = 0;
  bool keep = true;
  while (keep) {
    func_1d_orig2(QdotBsq, K, D, Bsq, Qtsq, x, dx, &f, &df, gamma);
    x[0] += dx[0];
    int inc = 0;
    while (((x[0] * x[0] * x[0] * (x[0] + 2. * Bsq) - QdotBsq * (2. * x[0] + Bsq)) <=
            x[0] * x[0] * (Qtsq - Bsq * Bsq)) &&
           (inc < 10)) {
      x[0] -= (1. * inc) * dx[0] / 10.;
      inc++;
    }
    errx = (x[0] == 0.) ? std::fabs(dx[0]) : std::fabs(dx[0] / x[0]);
    x[0] = std::fabs(x[0]);
    if ((std::fabs(errx) <= NEWT_TOL) && (doing_extra == 0) && (EXTRA_NEWT_ITER > 0)) doing_extra = 1;
    if (doing_extra == 1) i_extra++;
    if (((std::fabs(errx) <= NEWT_TOL) && (doing_extra == 0)) || (i_extra > EXTRA_NEWT_ITER) ||
        (n_iter >= (MAX_NEWT_ITER - 1)))
      keep = false;
    n_iter++;
  }
  *iters = n_iter;
  if (!std::isfinite(f) || !std::isfinite(df) || !std::isfinite(x[0])) return 2;
  if (std::fabs(errx) > MIN_NEWT_TOL) return 1;
  return 0;
}

// Scheme1D::invert (scheme_1d.h:411-568).
int c2p_1d(real U[NPRIM_RECV], real prim[NPRIM_RECV], real gamma, const real gcov[4][4],
           const real gcon[4][4], int *iters) {
  *iters = 0;
  const real K = prim[ENT];
  prim[BX1] = U[BX1];
  prim[BX2] = U[BX2];
  prim[BX3] = U[BX3];
  NobleInv v;
  noble_setup(U, gcov, gcon, v);
  real utsq;
  if (!noble_utsq(prim, gcov, &utsq)) return 2;
  const real uvsq = 1. + utsq;
  real Gamma = std::sqrt(uvsq);
  real rho0 = v.D / Gamma;
  real p = (K * std::pow(rho0, gamma));
  real uu = (p / (gamma - 1.));
  const real w = rho0 + uu + p;
  real W_last = w * uvsq;
  W_last = bump_w(W_last, v.Bsq, v.QdotBsq, v.Qtsq);
  real x[1] = {W_last};
  const int retval = gnr_1d(v.QdotBsq, K, v.D, v.Bsq, v.Qtsq, x, gamma, iters);
  const real W = x[0];
  if ((retval != 0) || (W == FAIL_VAL)) return retval * 100 + 1;
  if (W <= 0. || W > W_TOO_BIG) return 3;
  const real vsq = vsq_of_w(v.Bsq, v.QdotBsq, v.Qtsq, W);
  if (vsq >= 1.) return 4;
  const real gtmp = std::sqrt(1. - vsq);
  Gamma = 1. / gtmp;
  rho0 = v.D * gtmp;
  p = (K * std::pow(rho0, gamma));
  uu = (p / (gamma - 1.));
  if ((rho0 <= 0.) || (uu <= 0.)) return 5;
  prim[RHO] = rho0;
  prim[ENY] = uu;
  prim[ENT] = (gamma - 1.) * uu * std::pow(rho0, -gamma);
  noble_velocity(v, W, Gamma, prim);
  prim[BX1] = U[BX1];
  prim[BX2] = U[BX2];
  prim[BX3] = U[BX3];
  return 0;
}

// ------------------------------------------------------------ Kastaun 2021
// scheme_kastaun.h:84-108.
real kastaun_root(real mu, real D, real q, real bsq, real rsq, real rbsq, real gamma) {
  const real x = 1.0 / (1.0 + mu * bsq);
  const real xsq = x * x;
  const real rbarsq = x * (rsq * x + mu * (x + 1.0) * rbsq);
  const real musq = mu * mu;
  real vsq = musq * rbarsq;
  vsq = kmin(kmax(vsq, 0.0), 1.0 - 1.0e-14);
  const real il = std::sqrt(1.0 - vsq);
  const real qbar = (g_plant == PLANT_KASTAUN_QBAR)
                        ? q + 0.5 * bsq - 0.5 * musq * xsq * (bsq * rsq - rbsq)
                        : q - 0.5 * bsq - 0.5 * musq * xsq * (bsq * rsq - rbsq);
  const real eod = qbar - mu * rbarsq + 1.0;
  const real rho = D * il;
  real eps = D * eod - rho;
  eps = kmax(eps, 1.0e-300);
  const real pr = (gamma - 1.0) * eps;
  const real hh = 1.0 + (eps + pr) / rho;
  const real nua = hh * il;
  const real nub = eod + pr / D;
  const real nu = kmax(nua, nub);
  const real muh = 1.0 / (nu + mu * rbarsq);
  return mu - muh;
}

// Illinois false position on [0,1] (scheme_kastaun.h:110-165).
static int false_position(real *mu_out, real D, real q, real bsq, real rsq, real rbsq,
                          real gamma, int *iters) {
  real lo = 0.0, hi = 1.0;
  real flo = kastaun_root(lo, D, q, bsq, rsq, rbsq, gamma);
  real fhi = kastaun_root(hi, D, q, bsq, rsq, rbsq, gamma);
  *iters = 0;
  if (!std::isfinite(flo) || !std::isfinite(fhi)) return 1;
  if (std::fabs(flo) <= NEWT_TOL) { *mu_out = lo; return 0; }
  if (std::fabs(fhi) <= NEWT_TOL) { *mu_out = hi; return 0; }
  if (flo * fhi > 0.0) return 2;
  int side = 0;
  real mu = lo;
  for (int it = 0; it < MAX_NEWT_ITER; ++it) {
    const real old = mu;
    mu = (fhi * lo - flo * hi) / (fhi - flo);
    const real fm = kastaun_root(mu, D, q, bsq, rsq, rbsq, gamma);
    *iters = it + 1;
    if (!std::isfinite(mu) || !std::isfinite(fm)) return 3;
    if (std::fabs((mu - old) / (mu + 1.0e-300)) <= NEWT_TOL) { *mu_out = mu; return 0; }
    if (fm * flo >= 0.0) {
      if (side == 1) {
        const real wgt = 1.0 - fm / flo;
        fhi = (wgt > 0.0) ? fhi * wgt : 0.5 * fhi;
      }
      lo = mu;
      flo = fm;
      side = 1;
    } else {
      if (side == -1) {
        const real wgt = 1.0 - fm / fhi;
        flo = (wgt > 0.0) ? flo * wgt : 0.5 * flo;
      }
      hi = mu;
      fhi = fm;
      side = -1;
    }
  }
  return 4;
}

// SchemeKastaun::invert (scheme_kastaun.h:167-259).
int c2p_kastaun(real U[NPRIM_RECV], real prim[NPRIM_RECV], real gamma, const real gcov[4][4],
                const real gcon[4][4], int *iters) {
  *iters = 0;
  const real g[6] = {gcov[1][1], gcov[1][2], gcov[1][3], gcov[2][2], gcov[2][3], gcov[3][3]};
  const real det = g[0] * g[3] * g[5] + 2.0 * g[1] * g[2] * g[4] - g[0] * g[4] * g[4] -
                   g[1] * g[1] * g[5] - g[2] * g[2] * g[3];
  if (!(det > 0.0) || !std::isfinite(det)) return 1;
  const real idet = 1.0 / det;
  real gi[6];
  gi[0] = (g[3] * g[5] - g[4] * g[4]) * idet;
  gi[1] = (g[2] * g[4] - g[1] * g[5]) * idet;
  gi[2] = (g[1] * g[4] - g[2] * g[3]) * idet;
  gi[3] = (g[0] * g[5] - g[2] * g[2]) * idet;
  gi[4] = (g[1] * g[2] - g[0] * g[4]) * idet;
  gi[5] = (g[0] * g[3] - g[1] * g[1]) * idet;
  const real D = U[RHO];
  if (!(D > 0.0) || !std::isfinite(D)) return 2;
  real ncov[4], Qcov[4], Qcon[4];
  const real lapse = std::sqrt(-1.0 / gcon[0][0]);
  for (int i = 0; i < 4; ++i) ncov[i] = -lapse * (real)(i == 0);
  Qcov[0] = U[ENY];
  Qcov[1] = U[UX1];
  Qcov[2] = U[UX2];
  Qcov[3] = U[UX3];
  lower4(Qcov, gcon, Qcon);
  const real Qdotn = Qcon[0] * ncov[0];
  const real tau = -Qdotn - D;
  if (!std::isfinite(tau)) return 3;
  const real Sd[3] = {U[UX1], U[UX2], U[UX3]};
  real Su[3];
  Su[0] = gi[0] * Sd[0] + gi[1] * Sd[1] + gi[2] * Sd[2];
  Su[1] = gi[1] * Sd[0] + gi[3] * Sd[1] + gi[4] * Sd[2];
  Su[2] = gi[2] * Sd[0] + gi[4] * Sd[1] + gi[5] * Sd[2];
  const real Bu[3] = {U[BX1], U[BX2], U[BX3]};
  real Bd[3];
  Bd[0] = g[0] * Bu[0] + g[1] * Bu[1] + g[2] * Bu[2];
  Bd[1] = g[1] * Bu[0] + g[3] * Bu[1] + g[4] * Bu[2];
  Bd[2] = g[2] * Bu[0] + g[4] * Bu[1] + g[5] * Bu[2];
  const real Bsq = Bu[0] * Bd[0] + Bu[1] * Bd[1] + Bu[2] * Bd[2];
  const real Ssq = Su[0] * Sd[0] + Su[1] * Sd[1] + Su[2] * Sd[2];
  const real SdotB = Bu[0] * Sd[0] + Bu[1] * Sd[1] + Bu[2] * Sd[2];
  if (!std::isfinite(Bsq) || !std::isfinite(Ssq) || !std::isfinite(SdotB)) return 4;
  const real inv_D = 1.0 / D;
  const real sqrt_D = std::sqrt(D);
  const real bsq = Bsq * inv_D;
  const real rsq = Ssq * inv_D * inv_D;
  const real rbn = SdotB / (D * sqrt_D);
  const real rbsq = rbn * rbn;
  const real q = tau * inv_D;
  real mu = 0.0;
  const int rs = false_position(&mu, D, q, bsq, rsq, rbsq, gamma, iters);
  if (rs != 0) return 100 + rs;
  const real x = 1.0 / (1.0 + mu * bsq);
  const real rb = SdotB / (D * sqrt_D);
  const real rbarsq = x * (rsq * x + mu * (x + 1.0) * rbsq);
  const real vsq = mu * mu * rbarsq;
  if (!(vsq >= 0.0) || vsq >= 1.0 || !std::isfinite(vsq)) return 5;
  const real lorentz = 1.0 / std::sqrt(1.0 - vsq);
  const real rho = D / lorentz;
  const real qbar = q - 0.5 * bsq - 0.5 * mu * mu * x * x * (bsq * rsq - rbsq);
  const real eod = qbar - mu * rbarsq + 1.0;
  const real eps = D * eod - rho;
  if (!(rho > 0.0) || !(eps > 0.0) || !std::isfinite(eps)) return 6;
  const real ru[3] = {Su[0] * inv_D, Su[1] * inv_D, Su[2] * inv_D};
  const real bu[3] = {Bu[0] / sqrt_D, Bu[1] / sqrt_D, Bu[2] / sqrt_D};
  const real Wmux = lorentz * mu * x;
  prim[RHO] = rho;
  prim[ENY] = eps;
  prim[ENT] = (gamma - 1.) * eps * std::pow(rho, -gamma);
  for (int i = 0; i < 3; ++i) prim[UX1 + i] = Wmux * (ru[i] + rb * mu * bu[i]);
  prim[BX1] = U[BX1];
  prim[BX2] = U[BX2];
  prim[BX3] = U[BX3];
  return 0;
}

}  // namespace oracle
