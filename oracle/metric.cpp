// Oracle: coordinate maps, physical and code metrics, 4x4 algebra and the
// finite-difference Christoffel connection.  TEST INFRASTRUCTURE ONLY.
#include "oracle.h"

namespace oracle {

int g_plant = PLANT_NONE;

// MKS.h:15-23, BL.h:15-23 (identical maps), CKS.h:13-20, minkowski.h:13-20.
void physical_coords(int kind, const real x[4], real y[4], real h) {
  if (kind == MET_MKS || kind == MET_BL) {
    y[0] = x[0];
    y[1] = std::exp(x[1]);
    const real t1 = M_PI_2 * (x[2] + 1.0);
    const real t2 = 0.5 * h * std::sin(M_PI * (x[2] + 1.0));
    y[2] = t1 + t2;
    y[3] = x[3];
  } else {
    for (int n = 0; n < 4; ++n) y[n] = x[n];
  }
}

// Kerr-Schild in spherical-like coordinates (MKS.h:25-58).
static void metric_ks(const real x[4], real g[4][4], real h, real a) {
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) g[i][j] = 0.0;
  real y[4];
  physical_coords(MET_MKS, x, y, h);
  const real r = y[1];
  const real cth = std::cos(y[2]);
  const real sth = std::sin(y[2]);
  const real sigma = r * r + a * a * cth * cth;
  g[0][0] = -1.0 + 2.0 * r / sigma;
  g[0][1] = 2.0 * r / sigma;
  g[0][3] = -2.0 * a * r * sth * sth / sigma;
  g[1][0] = g[0][1];
  g[1][1] = 1.0 + 2.0 * r / sigma;
  g[1][3] = -a * (1.0 + 2.0 * r / sigma) * sth * sth;
  g[2][2] = sigma;
  g[3][0] = g[0][3];
  g[3][1] = g[1][3];
  g[3][3] = sth * sth * (sigma + a * a * sth * sth * (1.0 + 2.0 * r / sigma));
}

// Boyer-Lindquist (BL.h:25-58).
static void metric_bl(const real x[4], real g[4][4], real h, real a) {
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) g[i][j] = 0.0;
  real y[4];
  physical_coords(MET_BL, x, y, h);
  const real r = y[1];
  const real cth = std::cos(y[2]);
  const real sth = std::sin(y[2]);
  const real sth2 = sth * sth;
  const real r2 = r * r;
  const real a2 = a * a;
  const real mu = 1.0 + a2 * cth * cth / r2;
  const real delta = 1.0 - 2.0 / r + a2 / r2;
  const real inv_r_mu = 1.0 / (r * mu);
  g[0][0] = -1.0 + 2.0 * inv_r_mu;
  g[0][3] = -2.0 * a * sth2 * inv_r_mu;
  g[1][1] = mu / delta;
  // BL.h:54 writes r^2/mu here; Kerr's Sigma is r^2 mu (as in
  // fm_torus_cks/problem_generator.cpp:56).  The restatement keeps the
  // reference's form; the planted switch gives Kerr's.
  g[2][2] = (g_plant == PLANT_BL_GTHTH) ? r2 * mu : r2 / mu;
  g[3][0] = g[0][3];
  g[3][3] = r2 * sth2 * (1.0 + a2 / r2 + 2.0 * a2 * sth2 / (r * r2 * mu));
}

// Cartesian Kerr-Schild (CKS.h:22-69).
static void metric_cks(const real x[4], real g[4][4], real a) {
  const real xx = x[1], yy = x[2], zz = x[3];
  const real r2c = xx * xx + yy * yy + zz * zz;
  const real a2 = a * a;
  const real tmp = r2c - a2;
  const real sqrtarg = tmp * tmp + 4.0 * a2 * zz * zz;
  const real r = std::sqrt(0.5 * (tmp + std::sqrt(sqrtarg)));
  const real rs = (r > 1.0e-15) ? r : real(1.0e-15);
  const real r2 = rs * rs;
  const real r3 = r2 * rs;
  const real r4 = r2 * r2;
  const real xi_den = r4 + a2 * zz * zz;
  const real xi = (xi_den > 0.0) ? (2.0 * r3 / xi_den) : real(0.0);
  const real den_l = r2 + a2;
  const real l[4] = {1.0, (rs * xx + a * yy) / den_l, (rs * yy - a * xx) / den_l, zz / rs};
  for (int m = 0; m < 4; ++m)
    for (int n = 0; n < 4; ++n) g[m][n] = 0.0;
  g[0][0] = -1.0;
  g[1][1] = 1.0;
  g[2][2] = 1.0;
  g[3][3] = 1.0;
  for (int m = 0; m < 4; ++m)
    for (int n = 0; n < 4; ++n) g[m][n] += xi * l[m] * l[n];
}

void physical_metric(int kind, const real x[4], real g[4][4], real h, real a) {
  switch (kind) {
    case MET_MKS: metric_ks(x, g, h, a); break;
    case MET_BL: metric_bl(x, g, h, a); break;
    case MET_CKS: metric_cks(x, g, a); break;
    default:  // minkowski.h:22-31
      for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) g[i][j] = (real)((i == j) * (1 - 2 * (i == 0)));
  }
}

// Code-basis metric from a central-difference Jacobian with delta = 1e-5
// (MKS.h:60-107; identical in BL.h, CKS.h, minkowski.h).  The contraction
// keeps the reference's index order J[row][p] * J[q][col].
void code_metric(int kind, const real x[4], real g[4][4], real h, real a) {
  const real delta = 1.0e-5;
  real gp[4][4];
  physical_metric(kind, x, gp, h, a);
  real jac[4][4];
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) jac[r][c] = 0.0;
  for (int c = 0; c < 4; ++c) {
    real xh[4], xl[4], yh[4], yl[4];
    for (int n = 0; n < 4; ++n) { xh[n] = x[n]; xl[n] = x[n]; }
    xh[c] += delta;
    xl[c] -= delta;
    physical_coords(kind, xh, yh, h);
    physical_coords(kind, xl, yl, h);
    for (int r = 0; r < 4; ++r) jac[r][c] = (yh[r] - yl[r]) / (xh[c] - xl[c]);
  }
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) {
      real acc = 0.0;
      for (int p = 0; p < 4; ++p)
        for (int q = 0; q < 4; ++q) acc += gp[p][q] * jac[r][p] * jac[q][c];
      g[r][c] = acc;
    }
}

// Cofactor expansion along row 0 (tensor_algebra.h:32-62).
real determinant4(const real m[4][4]) {
  return m[0][0] * (m[1][1] * (m[2][2] * m[3][3] - m[2][3] * m[3][2]) -
                    m[1][2] * (m[2][1] * m[3][3] - m[2][3] * m[3][1]) +
                    m[1][3] * (m[2][1] * m[3][2] - m[2][2] * m[3][1])) -
         m[0][1] * (m[1][0] * (m[2][2] * m[3][3] - m[2][3] * m[3][2]) -
                    m[1][2] * (m[2][0] * m[3][3] - m[2][3] * m[3][0]) +
                    m[1][3] * (m[2][0] * m[3][2] - m[2][2] * m[3][0])) +
         m[0][2] * (m[1][0] * (m[2][1] * m[3][3] - m[2][3] * m[3][1]) -
                    m[1][1] * (m[2][0] * m[3][3] - m[2][3] * m[3][0]) +
                    m[1][3] * (m[2][0] * m[3][1] - m[2][1] * m[3][0])) -
         m[0][3] * (m[1][0] * (m[2][1] * m[3][2] - m[2][2] * m[3][1]) -
                    m[1][1] * (m[2][0] * m[3][2] - m[2][2] * m[3][0]) +
                    m[1][2] * (m[2][0] * m[3][1] - m[2][1] * m[3][0]));
}

// Gauss-Jordan on [M | I] swapping rows only on an exact zero pivot
// (tensor_algebra.h:64-124).  Returns 0 (and leaves inv untouched) when singular.
int invert4(const real m[4][4], real inv[4][4]) {
  real aug[4][8];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      aug[i][j] = m[i][j];
      aug[i][j + 4] = (i == j) ? 1.0 : 0.0;
    }
  for (int i = 0; i < 4; ++i) {
    if (aug[i][i] == 0.0) {
      int sw = -1;
      for (int j = i + 1; j < 4; ++j)
        if (aug[j][i] != 0.0) { sw = j; break; }
      if (sw == -1) return 0;
      for (int j = 0; j < 8; ++j) {
        const real t = aug[i][j];
        aug[i][j] = aug[sw][j];
        aug[sw][j] = t;
      }
    }
    const real piv = aug[i][i];
    for (int j = 0; j < 8; ++j) aug[i][j] /= piv;
    for (int j = 0; j < 4; ++j) {
      if (j == i) continue;
      const real f = aug[j][i];
      for (int k = 0; k < 8; ++k) aug[j][k] -= f * aug[i][k];
    }
  }
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) inv[i][j] = aug[i][j + 4];
  return 1;
}

// christoffel.h:53-62.
int metric_at(int kind, const real x[4], real a, real h, real gcov[4][4],
              real gcon[4][4], real *gdet) {
  code_metric(kind, x, gcov, h, a);
  const int ok = invert4(gcov, gcon);
  *gdet = determinant4(gcov);
  return ok;
}

// Gamma^mu_{alpha beta} by central differences (delta = 1e-5) of the code
// metric (christoffel.h:74-148).  Minkowski returns zeros (:90).
void christoffel(int kind, const real x[4], real a, real h, real conn[4][4][4]) {
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j)
      for (int k = 0; k < 4; ++k) conn[i][j][k] = 0.0;
  if (kind == MET_MINKOWSKI) return;
  real gcov[4][4], gcon[4][4], gdet;
  metric_at(kind, x, a, h, gcov, gcon, &gdet);
  const real delta = 1.0e-5;
  real dg[4][4][4];
  for (int d = 0; d < 4; ++d) {
    for (int r = 0; r < 4; ++r)
      for (int c = 0; c < 4; ++c) dg[d][r][c] = 0.0;
    if (d > 0) {
      real xp[4] = {x[0], x[1], x[2], x[3]};
      real xm[4] = {x[0], x[1], x[2], x[3]};
      xp[d] += delta;
      xm[d] -= delta;
      real gp[4][4], gm[4][4];
      code_metric(kind, xp, gp, h, a);
      code_metric(kind, xm, gm, h, a);
      for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) dg[d][r][c] = (gp[r][c] - gm[r][c]) / (2.0 * delta);
    }
  }
  real cc[4][4][4];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j)
      for (int k = 0; k < 4; ++k)
        cc[i][j][k] = (g_plant == PLANT_CONN_INDEX)
                          ? 0.5 * (dg[i][j][k] + dg[k][i][j] - dg[j][i][k])
                          : 0.5 * (dg[j][i][k] + dg[k][i][j] - dg[i][j][k]);
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j)
      for (int k = 0; k < 4; ++k) {
        real v = 0.0;
        for (int l = 0; l < 4; ++l) v += gcon[i][l] * cc[l][j][k];
        conn[i][j][k] = v;
      }
}

}  // namespace oracle
