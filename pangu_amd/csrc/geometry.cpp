// Host-side geometry tables for the device kernels.
//
// PANGU recomputes the code-basis metric (a central-difference Jacobian of
// the coordinate map, MKS.h:60-107), its Gauss-Jordan inverse
// (tensor_algebra.h:64-124), determinant and the finite-difference
// Christoffel symbols (christoffel.h:74-148) inside every kernel, ~15 times
// per zone per stage.  For Minkowski/BL/MKS the metric does not depend on x3,
// so one 2D (j,i) table per distinct block footprint serves every block and
// every k-plane; CKS gets 3D tables.  Values are computed once here, in FP64
// with the reference's operation order (build with -ffp-contract=off), so
// device kernels read exactly the numbers a recomputation would give.
#include <cmath>
#include <cstring>
#include <vector>

#include "common.h"
#include "host.h"

namespace pangu {
namespace geom {

typedef double M4[4][4];

// Coordinate maps: MKS and BL share r = e^{x1},
// theta = pi/2 (x2+1) + h/2 sin(pi (x2+1)) (MKS.h:15-23); CKS and Minkowski
// are the identity.
static void to_physical(int kind, const double x[4], double y[4], double h) {
  if (kind == PANGU_METRIC_MKS || kind == PANGU_METRIC_BL) {
    const double th_lin = M_PI_2 * (x[2] + 1.0);
    const double th_mod = 0.5 * h * std::sin(M_PI * (x[2] + 1.0));
    y[0] = x[0];
    y[1] = std::exp(x[1]);
    y[2] = th_lin + th_mod;
    y[3] = x[3];
    return;
  }
  y[0] = x[0];
  y[1] = x[1];
  y[2] = x[2];
  y[3] = x[3];
}

static void zero(M4 g) {
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) g[r][c] = 0.0;
}

// Kerr-Schild in (t, r, theta, phi) (MKS.h:25-58).
static void kerr_schild(const double x[4], M4 g, double h, double a) {
  zero(g);
  double y[4];
  to_physical(PANGU_METRIC_MKS, x, y, h);
  const double r = y[1], c = std::cos(y[2]), s = std::sin(y[2]);
  const double sig = r * r + a * a * c * c;
  const double two_r_sig = 2.0 * r / sig;
  g[0][0] = -1.0 + two_r_sig;
  g[0][1] = two_r_sig;
  g[0][3] = -2.0 * a * r * s * s / sig;
  g[1][0] = g[0][1];
  g[1][1] = 1.0 + two_r_sig;
  g[1][3] = -a * (1.0 + two_r_sig) * s * s;
  g[2][2] = sig;
  g[3][0] = g[0][3];
  g[3][1] = g[1][3];
  g[3][3] = s * s * (sig + a * a * s * s * (1.0 + two_r_sig));
}

// Boyer-Lindquist (BL.h:25-58).
static void boyer_lindquist(const double x[4], M4 g, double h, double a) {
  zero(g);
  double y[4];
  to_physical(PANGU_METRIC_BL, x, y, h);
  const double r = y[1], c = std::cos(y[2]), s = std::sin(y[2]);
  const double s2 = s * s, r2 = r * r, a2 = a * a;
  const double mu = 1.0 + a2 * c * c / r2;
  const double del = 1.0 - 2.0 / r + a2 / r2;
  const double irm = 1.0 / (r * mu);
  g[0][0] = -1.0 + 2.0 * irm;
  g[0][3] = -2.0 * a * s2 * irm;
  g[1][1] = mu / del;
  g[2][2] = r2 / mu;
  g[3][0] = g[0][3];
  g[3][3] = r2 * s2 * (1.0 + a2 / r2 + 2.0 * a2 * s2 / (r * r2 * mu));
}

// Cartesian Kerr-Schild g = eta + xi l l (CKS.h:22-69).
static void cartesian_ks(const double x[4], M4 g, double a) {
  const double X = x[1], Y = x[2], Z = x[3];
  const double a2 = a * a;
  const double q = (X * X + Y * Y + Z * Z) - a2;
  const double r = std::sqrt(0.5 * (q + std::sqrt(q * q + 4.0 * a2 * Z * Z)));
  const double rr = (r > 1.0e-15) ? r : 1.0e-15;
  const double r2 = rr * rr, r3 = r2 * rr, r4 = r2 * r2;
  const double den = r4 + a2 * Z * Z;
  const double xi = (den > 0.0) ? (2.0 * r3 / den) : 0.0;
  const double dl = r2 + a2;
  const double l[4] = {1.0, (rr * X + a * Y) / dl, (rr * Y - a * X) / dl, Z / rr};
  zero(g);
  g[0][0] = -1.0;
  g[1][1] = 1.0;
  g[2][2] = 1.0;
  g[3][3] = 1.0;
  for (int m = 0; m < 4; ++m)
    for (int n = 0; n < 4; ++n) g[m][n] += xi * l[m] * l[n];
}

static void physical(int kind, const double x[4], M4 g, double h, double a) {
  if (kind == PANGU_METRIC_MKS) return kerr_schild(x, g, h, a);
  if (kind == PANGU_METRIC_BL) return boyer_lindquist(x, g, h, a);
  if (kind == PANGU_METRIC_CKS) return cartesian_ks(x, g, a);
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) g[r][c] = (r == c) ? (r == 0 ? -1.0 : 1.0) : 0.0;
}

// g_code[r][c] = sum_p sum_q g[p][q] J[r][p] J[q][c] with a central
// difference Jacobian, delta = 1e-5 (MKS.h:60-107).
void code_metric(int kind, const double x[4], M4 out, double h, double a) {
  M4 g;
  physical(kind, x, g, h, a);
  M4 J;
  zero(J);
  const double del = 1.0e-5;
  for (int c = 0; c < 4; ++c) {
    double xp[4], xm[4], yp[4], ym[4];
    std::memcpy(xp, x, sizeof(xp));
    std::memcpy(xm, x, sizeof(xm));
    xp[c] += del;
    xm[c] -= del;
    to_physical(kind, xp, yp, h);
    to_physical(kind, xm, ym, h);
    const double span = xp[c] - xm[c];
    for (int r = 0; r < 4; ++r) J[r][c] = (yp[r] - ym[r]) / span;
  }
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) {
      double s = 0.0;
      for (int p = 0; p < 4; ++p)
        for (int q = 0; q < 4; ++q) s += g[p][q] * J[r][p] * J[q][c];
      out[r][c] = s;
    }
}

// Row-0 cofactor determinant (tensor_algebra.h:32-62).
double det4(const M4 m) {
  const double c00 = m[2][2] * m[3][3] - m[2][3] * m[3][2];
  const double c01 = m[2][1] * m[3][3] - m[2][3] * m[3][1];
  const double c02 = m[2][1] * m[3][2] - m[2][2] * m[3][1];
  const double c10 = m[2][0] * m[3][3] - m[2][3] * m[3][0];
  const double c12 = m[2][0] * m[3][2] - m[2][2] * m[3][0];
  const double c20 = m[2][0] * m[3][1] - m[2][1] * m[3][0];
  return m[0][0] * (m[1][1] * c00 - m[1][2] * c01 + m[1][3] * c02) -
         m[0][1] * (m[1][0] * c00 - m[1][2] * c10 + m[1][3] * c12) +
         m[0][2] * (m[1][0] * c01 - m[1][1] * c10 + m[1][3] * c20) -
         m[0][3] * (m[1][0] * c02 - m[1][1] * c12 + m[1][2] * c20);
}

// Gauss-Jordan on [m | I], row swap only on an exact zero pivot
// (tensor_algebra.h:64-124).  Returns false when singular.
bool inv4(const M4 m, M4 out) {
  double A[4][8];
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) {
      A[r][c] = m[r][c];
      A[r][c + 4] = (r == c) ? 1.0 : 0.0;
    }
  for (int p = 0; p < 4; ++p) {
    if (A[p][p] == 0.0) {
      int sw = -1;
      for (int r = p + 1; r < 4 && sw < 0; ++r)
        if (A[r][p] != 0.0) sw = r;
      if (sw < 0) return false;
      for (int c = 0; c < 8; ++c) std::swap(A[p][c], A[sw][c]);
    }
    const double piv = A[p][p];
    for (int c = 0; c < 8; ++c) A[p][c] /= piv;
    for (int r = 0; r < 4; ++r) {
      if (r == p) continue;
      const double f = A[r][p];
      for (int c = 0; c < 8; ++c) A[r][c] -= f * A[p][c];
    }
  }
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) out[r][c] = A[r][c + 4];
  return true;
}

// Gamma^i_{jk} = g^{il} Gamma_{ljk}, Gamma_{ljk} = (d_j g_lk + d_k g_lj - d_l g_jk)/2,
// derivatives by central differences of the code metric (christoffel.h:74-148).
void connection(int kind, const double x[4], const M4 gcon, double h, double a,
                double conn[64]) {
  for (int n = 0; n < 64; ++n) conn[n] = 0.0;
  if (kind == PANGU_METRIC_MINKOWSKI) return;
  double dg[4][4][4];
  const double del = 1.0e-5;
  for (int d = 0; d < 4; ++d) {
    for (int r = 0; r < 4; ++r)
      for (int c = 0; c < 4; ++c) dg[d][r][c] = 0.0;
    if (d == 0) continue;
    double xp[4], xm[4];
    std::memcpy(xp, x, sizeof(xp));
    std::memcpy(xm, x, sizeof(xm));
    xp[d] += del;
    xm[d] -= del;
    M4 gp, gm;
    code_metric(kind, xp, gp, h, a);
    code_metric(kind, xm, gm, h, a);
    for (int r = 0; r < 4; ++r)
      for (int c = 0; c < 4; ++c) dg[d][r][c] = (gp[r][c] - gm[r][c]) / (2.0 * del);
  }
  double low[4][4][4];
  for (int l = 0; l < 4; ++l)
    for (int j = 0; j < 4; ++j)
      for (int k = 0; k < 4; ++k) low[l][j][k] = 0.5 * (dg[j][l][k] + dg[k][l][j] - dg[l][j][k]);
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j)
      for (int k = 0; k < 4; ++k) {
        double s = 0.0;
        for (int l = 0; l < 4; ++l) s += gcon[i][l] * low[l][j][k];
        conn[16 * i + 4 * j + k] = s;
      }
}

// r used by the floors (primitive_fixer.cc:79-86).
static double floor_radius(int kind, const double x[4], double a) {
  if (kind == PANGU_METRIC_MKS) return std::exp(x[1]);
  if (kind == PANGU_METRIC_CKS) {
    const double a2 = a * a;
    const double rad2 = x[1] * x[1] + x[2] * x[2] + x[3] * x[3];
    const double sa = (rad2 - a2) * (rad2 - a2) + 4.0 * a2 * x[3] * x[3];
    return std::sqrt(0.5 * (rad2 - a2 + std::sqrt(sa)));
  }
  return 1.0;
}

void fill_centre(const pangu_config &cfg, const double x[4], double *out, int64_t stride) {
  M4 gcov, gcon;
  code_metric(cfg.metric, x, gcov, cfg.h, cfg.a);
  zero(gcon);
  inv4(gcov, gcon);
  const double gd = det4(gcov);
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) {
      out[(CG_GCOV + 4 * r + c) * stride] = gcov[r][c];
      out[(CG_GCON + 4 * r + c) * stride] = gcon[r][c];
    }
  out[CG_SG * stride] = std::sqrt(std::fabs(gd));
  out[CG_ALPHA * stride] = 1.0 / std::sqrt(-gcon[0][0]);
  out[CG_LAPSE * stride] = std::sqrt(-1.0 / gcon[0][0]);
  const double r = floor_radius(cfg.metric, x, cfg.a);
  out[CG_RHOFL * stride] = cfg.density_floor * std::pow(std::fmax(r, 1.0), cfg.density_floor_pow);
  out[CG_ENGFL * stride] = cfg.energy_floor * std::pow(std::fmax(r, 1.0), cfg.energy_floor_pow);
  out[CG_RAD * stride] = r;
  double conn[64];
  connection(cfg.metric, x, gcon, cfg.h, cfg.a, conn);
  for (int n = 0; n < 64; ++n) out[(CG_CONN + n) * stride] = conn[n];
}

void fill_metric(const pangu_config &cfg, const double x[4], double out[35]) {
  M4 gcov, gcon;
  code_metric(cfg.metric, x, gcov, cfg.h, cfg.a);
  zero(gcon);
  inv4(gcov, gcon);
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) {
      out[4 * r + c] = gcov[r][c];
      out[16 + 4 * r + c] = gcon[r][c];
    }
  out[32] = std::sqrt(std::fabs(det4(gcov)));
  out[33] = 1.0 / std::sqrt(-gcon[0][0]);
  out[34] = std::sqrt(-1.0 / gcon[0][0]);
}

void fill_metric_points(const pangu_config &cfg, const double *x12, int n, double *out) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    const double x[4] = {0.0, x12[2 * i], x12[2 * i + 1], 0.0};
    fill_metric(cfg, x, out + (int64_t)i * 35);
  }
}

void fill_face(const pangu_config &cfg, int d, const double x[4], double *out, int64_t stride) {
  M4 gcov, gcon;
  code_metric(cfg.metric, x, gcov, cfg.h, cfg.a);
  zero(gcon);
  const bool ok = inv4(gcov, gcon);
  const double gd = det4(gcov);
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) out[(FG_GCOV + 4 * r + c) * stride] = gcov[r][c];
  for (int c = 0; c < 4; ++c) out[(FG_GCON0 + c) * stride] = gcon[0][c];
  out[FG_GCONDD * stride] = gcon[d + 1][d + 1];
  out[FG_SG * stride] = std::sqrt(std::fabs(gd));
  out[FG_ALPHA * stride] = ok ? 1.0 / std::sqrt(-gcon[0][0]) : 0.0;
  out[FG_VALID * stride] = ok ? 1.0 : 0.0;
}

// Whole tables for one geometry set.  Centre entries at (Xc0(i), Xc1(j), x3)
// with x3 = Xc2(k) for 3D tables or the block's first interior plane
// otherwise; face entries of direction d at the lower face Xf_d.
void fill_centre_table(const pangu_config &cfg, const HostBlock &H, int ni, int nj, int nk,
                       bool has3d, double *out) {
  const int64_t n = (int64_t)ni * nj * nk;
  const double x3c = H.xc(2, H.s[2]);
#pragma omp parallel for schedule(dynamic, 1) collapse(2)
  for (int kk = 0; kk < nk; ++kk)
    for (int j = 0; j < nj; ++j)
      for (int i = 0; i < ni; ++i) {
        const double x[4] = {0.0, H.xc(0, i), H.xc(1, j), has3d ? H.xc(2, kk) : x3c};
        fill_centre(cfg, x, out + ((int64_t)kk * nj + j) * ni + i, n);
      }
}

void fill_face_table(const pangu_config &cfg, const HostBlock &H, int d, int fi, int fj, int fk,
                     bool has3d, double *out) {
  const int64_t n = (int64_t)fi * fj * fk;
  const double x3c = H.xc(2, H.s[2]);
#pragma omp parallel for schedule(dynamic, 1) collapse(2)
  for (int kk = 0; kk < fk; ++kk)
    for (int j = 0; j < fj; ++j)
      for (int i = 0; i < fi; ++i) {
        double x[4] = {0.0, H.xc(0, i), H.xc(1, j), has3d ? H.xc(2, kk) : x3c};
        if (d == 0) x[1] = H.xf(0, i);
        if (d == 1) x[2] = H.xf(1, j);
        if (d == 2) x[3] = H.xf(2, kk);
        fill_face(cfg, d, x, out + ((int64_t)kk * fj + j) * fi + i, n);
      }
}

}  // namespace geom
}  // namespace pangu
