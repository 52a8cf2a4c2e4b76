// Oracle: initial data.  TEST INFRASTRUCTURE ONLY.
//   problem 1: relativistic Orszag-Tang vortex on [0,1]^2 (Minkowski):
//     rho = 25/(36 pi), p = 5/(12 pi), v = 0.5(-sin 2 pi y, sin 2 pi x),
//     B = (-sin 2 pi y, sin 4 pi x)/sqrt(4 pi).  (Not the reference's OT
//     setup, which uses [-pi,pi]^2; this is the unit-box variant.)
//   problem 2: Fishbone-Moncrief torus in MKS with a SANE vector potential
//     (pangu/problem/fm_torus_mks/problem_generator.cpp:22-558); the random
//     pressure perturbation uses a counter-based generator keyed by
//     (seed, global cell index) instead of a thread-order-dependent pool.
#include <cfloat>
#include <cstring>

#include "oracle.h"
#include "oracle_sim.h"

namespace oracle {

#ifdef ORACLE_COUNTING
// The initial data are computed in double; in the counting build `real` is a
// counting type, so these adapters convert at the calls (init is not counted).
static void physical_coords(int kind, const double x[4], double y[4], double h) {
  real xr[4] = {x[0], x[1], x[2], x[3]}, yr[4];
  oracle::physical_coords(kind, xr, yr, h);
  for (int n = 0; n < 4; ++n) y[n] = yr[n];
}
static void code_metric(int kind, const double x[4], double g[4][4], double h, double a) {
  real xr[4] = {x[0], x[1], x[2], x[3]}, gr[4][4];
  oracle::code_metric(kind, xr, gr, h, a);
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) g[r][c] = gr[r][c];
}
static int invert4(const double m[4][4], double inv[4][4]) {
  real mr[4][4], ir[4][4];
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) { mr[r][c] = m[r][c]; ir[r][c] = inv[r][c]; }
  const int ok = oracle::invert4(mr, ir);
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) inv[r][c] = ir[r][c];
  return ok;
}
static double determinant4(const double m[4][4]) {
  real mr[4][4];
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) mr[r][c] = m[r][c];
  return oracle::determinant4(mr);
}
static void calc_state(const real prim[NPRIM], const double gcov[4][4], const double gcon[4][4], State &s) {
  real gc[4][4], gn[4][4];
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) { gc[r][c] = gcov[r][c]; gn[r][c] = gcon[r][c]; }
  oracle::calc_state(prim, gc, gn, s);
}
#endif

// splitmix64 finaliser of (seed, cell id) -> uniform double in [-1, 1).
static double keyed_uniform(uint64_t seed, uint64_t id) {
  uint64_t z = id + seed * 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  z = z ^ (z >> 31);
  const double u = (double)(z >> 11) * (1.0 / 9007199254740992.0);
  return -1.0 + 2.0 * u;
}

// fm_torus_mks/problem_generator.cpp:22-35.
static real lfish_calc(real r, real a) {
  return (((std::pow(a, 2.) - 2. * a * std::sqrt(r) + std::pow(r, 2.)) *
           ((-2. * a * r * (std::pow(a, 2.) - 2. * a * std::sqrt(r) + std::pow(r, 2.))) /
                std::sqrt(2. * a * std::sqrt(r) + (-3. + r) * r) +
            ((a + (-2. + r) * std::sqrt(r)) * (std::pow(r, 3.) + std::pow(a, 2.) * (2. + r))) /
                std::sqrt(1 + (2. * a) / std::pow(r, 1.5) - 3. / r))) /
          (std::pow(r, 3.) * std::sqrt(2. * a * std::sqrt(r) + (-3. + r) * r) *
           (std::pow(a, 2.) + (-2. + r) * r)));
}

// :37-48.
static real solve_ut(const real g[4][4], real u1, real u2, real u3) {
  const real AA = g[0][0];
  const real BB = 2.0 * (g[0][1] * u1 + g[0][2] * u2 + g[0][3] * u3);
  const real CC = 1.0 + g[1][1] * u1 * u1 + g[2][2] * u2 * u2 + g[3][3] * u3 * u3 +
                  2.0 * (g[1][2] * u1 * u2 + g[1][3] * u1 * u3 + g[2][3] * u2 * u3);
  const real discr = BB * BB - 4.0 * AA * CC;
  return (-BB - std::sqrt(discr)) / (2.0 * AA);
}

// :50-101.
static void bl_to_code_ucon(const real x[4], real h, real a, const real ubl[4], real uc[4]) {
  real y[4];
  physical_coords(MET_BL, x, y, h);
  const real r = y[1];
  real tr[4][4], tmp[4], dxdxp[4][4], dxpdx[4][4];
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) tr[i][j] = (real)(i == j);
  tr[0][1] = 2.0 * r / (r * r - 2.0 * r + a * a);
  tr[3][1] = a / (r * r - 2.0 * r + a * a);
  for (int i = 0; i < 4; i++) tmp[i] = 0.0;
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) tmp[i] += tr[i][j] * ubl[j];
  for (int c = 0; c < 4; c++) {
    real xh[4], xl[4], yh[4], yl[4];
    for (int n = 0; n < 4; ++n) { xh[n] = x[n]; xl[n] = x[n]; }
    xh[c] += 1e-5;
    xl[c] -= 1e-5;
    physical_coords(MET_BL, xh, yh, h);
    physical_coords(MET_BL, xl, yl, h);
    for (int r2 = 0; r2 < 4; r2++) dxdxp[r2][c] = (yh[r2] - yl[r2]) / (xh[c] - xl[c]);
  }
  invert4(dxdxp, dxpdx);
  for (int i = 0; i < 4; i++) {
    uc[i] = 0.0;
    for (int j = 0; j < 4; j++) uc[i] += dxpdx[i][j] * tmp[j];
  }
}

#ifdef ORACLE_COUNTING
static double solve_ut(const double g[4][4], double u1, double u2, double u3) {
  real gr[4][4];
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) gr[r][c] = g[r][c];
  return solve_ut(gr, real(u1), real(u2), real(u3));
}
static void bl_to_code_ucon(const double x[4], double h, double a, const double ubl[4], double uc[4]) {
  real xr[4], ur[4], cr[4];
  for (int n = 0; n < 4; ++n) { xr[n] = x[n]; ur[n] = ubl[n]; }
  bl_to_code_ucon(xr, real(h), real(a), ur, cr);
  for (int n = 0; n < 4; ++n) uc[n] = cr[n];
}
#endif

static void init_ot(Sim &S) {
  const double gam = S.p.gamma;
  for (int b = 0; b < S.nblk; ++b) {
    const Block &B = S.blk[b];
    for (int k = 0; k < B.n[2]; ++k)
      for (int j = 0; j < B.n[1]; ++j)
        for (int i = 0; i < B.n[0]; ++i) {
          const int64_t id = B.idx(k, j, i);
          const double x = Xc(B, 0, i), y = Xc(B, 1, j);
          const double pi = M_PI;
          const double rho = 25.0 / (36.0 * pi);
          const double pr = 5.0 / (12.0 * pi);
          const double u = pr / (gam - 1.0);
          const double vx = -0.5 * std::sin(2.0 * pi * y);
          const double vy = 0.5 * std::sin(2.0 * pi * x);
          const double v2 = vx * vx + vy * vy;
          const double W = 1.0 / std::sqrt(1.0 - v2);
          const double b0 = 1.0 / std::sqrt(4.0 * pi);
          S.Pc(0, b, 0)[id] = rho;
          S.Pc(0, b, 1)[id] = u;
          S.Pc(0, b, 2)[id] = W * vx;
          S.Pc(0, b, 3)[id] = W * vy;
          S.Pc(0, b, 4)[id] = 0.0;
          if (S.p.enable_B) {
            S.Pc(0, b, S.iBX)[id] = -b0 * std::sin(2.0 * pi * y);
            S.Pc(0, b, S.iBX + 1)[id] = b0 * std::sin(4.0 * pi * x);
            S.Pc(0, b, S.iBX + 2)[id] = 0.0;
          }
          const double ent = (gam - 1.0) * u * std::pow(rho, -gam);
          S.Pc(0, b, S.iENT)[id] = ent;
          if (S.iKEL >= 0) S.Pc(0, b, S.iKEL)[id] = S.cfg.fel_0 * ent;
        }
  }
}

static int init_torus(Sim &S) {
  const oracle_cfg &c = S.cfg;
  const double gam = S.p.gamma, a = S.p.a, h = S.p.h;
  const double rin = c.rin, kappa = c.kappa, pert = c.perturbation;
  const double a2 = a * a;
  const double l = lfish_calc(c.rmax, a);
  const double thin = M_PI_2;
  const double sthin = std::sin(thin), cthin = std::cos(thin);
  const double DDin = rin * rin - 2. * rin + a * a;
  const double AAin = (rin * rin + a * a) * (rin * rin + a * a) - DDin * a * a * sthin * sthin;
  const double SSin = rin * rin + a * a * cthin * cthin;
  // global extended index for the keyed generator
  int G[3], NX[3];
  for (int d = 0; d < 3; ++d) {
    G[d] = S.blk[0].ng[d];
    NX[d] = c.nx[d] + 2 * G[d];
  }
  // problem_generator.cpp:161-270 (entire block, ghosts included)
  for (int b = 0; b < S.nblk; ++b) {
    const Block &B = S.blk[b];
#pragma omp parallel for schedule(dynamic, 1)
    for (int k = 0; k < B.n[2]; ++k)
      for (int j = 0; j < B.n[1]; ++j)
        for (int i = 0; i < B.n[0]; ++i) {
          const int64_t id = B.idx(k, j, i);
          const int64_t gi = B.gofs[0] + i, gj = B.gofs[1] + j, gk = B.gofs[2] + k;
          const uint64_t gid = (uint64_t)((gk * NX[1] + gj) * NX[0] + gi);
          const double rnd = keyed_uniform(c.seed, gid);
          const double x1 = Xc(B, 0, i), x2 = Xc(B, 1, j), x3 = Xc(B, 2, k);
          const double xc[4] = {0.0, x1, x2, x3};
          double y[4];
          physical_coords(MET_MKS, xc, y, h);
          const double r = y[1], th = y[2];
          const double sth = std::sin(th), cth = std::cos(th);
          double gcov[4][4], gbl[4][4], gcon[4][4];
          code_metric(MET_MKS, xc, gcov, h, a);
          code_metric(MET_BL, xc, gbl, h, a);
          invert4(gcov, gcon);
          const double alpha = 1.0 / std::sqrt(-gcon[0][0]);
          const double beta1 = gcon[0][1] * alpha * alpha;
          const double beta2 = gcon[0][2] * alpha * alpha;
          const double beta3 = gcon[0][3] * alpha * alpha;
          double rho = 0.0, eint = 0.0, ent = 0.0, w1 = 0.0, w2 = 0.0, w3 = 0.0;
          const double DD = r * r - 2.0 * r + a2;
          const double AA = (r * r + a2) * (r * r + a2) - DD * a2 * sth * sth;
          const double SS = r * r + a2 * cth * cth;
          double lnh = 1.0;
          if (r >= rin) {
            lnh = 0.5 * std::log((1. + std::sqrt(1. + 4. * (l * l * SS * SS) * DD /
                                                         (AA * sth * AA * sth))) /
                                 (SS * DD / AA)) -
                  0.5 * std::sqrt(1. + 4. * (l * l * SS * SS) * DD / (AA * AA * sth * sth)) -
                  2. * a * r * l / AA -
                  (0.5 * std::log((1. + std::sqrt(1. + 4. * (l * l * SSin * SSin) * DDin /
                                                           (AAin * AAin * sthin * sthin))) /
                                  (SSin * DDin / AAin)) -
                   0.5 * std::sqrt(1. + 4. * (l * l * SSin * SSin) * DDin /
                                            (AAin * AAin * sthin * sthin)) -
                   2. * a * rin * l / AAin);
          }
          if (lnh >= 0.0 && r >= rin) {
            const double hm1 = std::exp(lnh) - 1.0;
            if (hm1 > 0.0) {
              rho = std::pow(hm1 * (gam - 1.0) / (kappa * gam), 1.0 / (gam - 1.0));
              eint = kappa * std::pow(rho, gam) / (gam - 1.0);
              eint *= (1.0 + pert * rnd);
              ent = (gam - 1) * eint * std::pow(rho, -gam);
              const double e2c = SS * SS * DD / (AA * AA * sth * sth);
              const double up1 = std::sqrt((-1.0 + std::sqrt(1.0 + 4.0 * l * l * e2c)) / 2.0);
              const double upbl = 2.0 * a * r * std::sqrt(1.0 + up1 * up1) / std::sqrt(AA * SS * DD) +
                                  std::sqrt(SS / AA) * up1 / sth;
              const double u0bl = solve_ut(gbl, 0.0, 0.0, upbl);
              const double ubl[4] = {u0bl, 0.0, 0.0, upbl};
              double uc[4];
              bl_to_code_ucon(xc, h, a, ubl, uc);
              w1 = uc[1] + beta1 * uc[0];
              w2 = uc[2] + beta2 * uc[0];
              w3 = uc[3] + beta3 * uc[0];
            }
          }
          S.Pc(0, b, 0)[id] = rho;
          S.Pc(0, b, 1)[id] = eint;
          S.Pc(0, b, 2)[id] = w1;
          S.Pc(0, b, 3)[id] = w2;
          S.Pc(0, b, 4)[id] = w3;
          if (S.p.enable_B) {
            S.Pc(0, b, S.iBX)[id] = 0.0;
            S.Pc(0, b, S.iBX + 1)[id] = 0.0;
            S.Pc(0, b, S.iBX + 2)[id] = 0.0;
          }
          S.Pc(0, b, S.iENT)[id] = ent;
          if (S.iKEL >= 0) S.Pc(0, b, S.iKEL)[id] = c.fel_0 * ent;
        }
  }
  // MeshPostInitialization (:291-558): maxima over entire blocks
  double rhomax = 0.0, umax = 0.0;
  for (int b = 0; b < S.nblk; ++b) {
    const Block &B = S.blk[b];
    double bm = 0.0, um = 0.0;
    for (int64_t id = 0; id < B.cells; ++id) {
      bm = kmax(bm, S.Pc(0, b, 0)[id]);
      um = kmax(um, S.Pc(0, b, 1)[id]);
    }
    rhomax = kmax(rhomax, bm);
    umax = kmax(umax, um);
  }
  // No cell inside the torus (unbound or unresolved): the reference divides
  // by the zero maximum (NaN everywhere); the zero state is kept instead
  // (floored to the atmosphere by the first FixPrimitive), with no field.
  if (!(rhomax > 0.0)) return 0;
  for (int b = 0; b < S.nblk; ++b) {
    const Block &B = S.blk[b];
    for (int64_t id = 0; id < B.cells; ++id) {
      S.Pc(0, b, 0)[id] /= rhomax;
      S.Pc(0, b, 1)[id] /= rhomax;
    }
    if (!S.p.enable_B) continue;
    // corner vector potential A_phi = max(<rho> - cut, 0) (:399-429, SANE)
    std::vector<double> A((size_t)B.cells, 0.0);
    for (int k = 0; k < B.n[2]; ++k)
      for (int j = 1; j < B.n[1]; ++j)
        for (int i = 1; i < B.n[0]; ++i) {
          const double *rho = S.Pc(0, b, 0);
          const double avg = 0.25 * (rho[B.idx(k, j, i)] + rho[B.idx(k, j, i - 1)] +
                                     rho[B.idx(k, j - 1, i)] + rho[B.idx(k, j - 1, i - 1)]);
          const double ex = avg - c.aphi_cut;
          A[B.idx(k, j, i)] = (ex > 0.0) ? ex : 0.0;
        }
    // B from curl A (:431-458)
    for (int k = B.s[2]; k <= B.e[2]; ++k)
      for (int j = B.s[1]; j <= B.e[1]; ++j)
        for (int i = B.s[0]; i <= B.e[0]; ++i) {
          const double xc[4] = {0.0, Xc(B, 0, i), Xc(B, 1, j), Xc(B, 2, k)};
          double gcov[4][4], gcon[4][4];
          code_metric(MET_MKS, xc, gcov, h, a);
          invert4(gcov, gcon);
          const double gdet = determinant4(gcov);
          const double sg = std::sqrt(std::fabs(gdet));
          const double dx1 = B.dx[0], dx2 = B.dx[1];
          const double a00 = A[B.idx(k, j - 1, i - 1)], a01 = A[B.idx(k, j, i - 1)];
          const double a10 = A[B.idx(k, j - 1, i)], a11 = A[B.idx(k, j, i)];
          S.Pc(0, b, S.iBX)[B.idx(k, j, i)] = -(a00 - a01 + a10 - a11) / (2.0 * dx2 * sg);
          S.Pc(0, b, S.iBX + 1)[B.idx(k, j, i)] = (a00 + a01 - a10 - a11) / (2.0 * dx1 * sg);
        }
  }
  if (!S.p.enable_B) return 0;
  // beta normalisation (:464-551)
  double bsqmax = 0.0;
  for (int b = 0; b < S.nblk; ++b) {
    const Block &B = S.blk[b];
    double bm = 0.0;
    for (int k = B.s[2]; k <= B.e[2]; ++k)
      for (int j = B.s[1]; j <= B.e[1]; ++j)
        for (int i = B.s[0]; i <= B.e[0]; ++i) {
          const double xc[4] = {0.0, Xc(B, 0, i), Xc(B, 1, j), Xc(B, 2, k)};
          double gcov[4][4], gcon[4][4];
          code_metric(MET_MKS, xc, gcov, h, a);
          invert4(gcov, gcon);
          real pc[NPRIM];
          S.load_prim(0, b, B.idx(k, j, i), pc);
          State st;
          calc_state(pc, gcov, gcon, st);
          bm = kmax(bm, st.bsq);
        }
    bsqmax = kmax(bsqmax, bm);
  }
  const double beta_min = 2.0 * (gam - 1.0) * (umax / rhomax) / (bsqmax + 1.0e-30);
  const double norm = std::sqrt(beta_min / c.beta);
  for (int b = 0; b < S.nblk; ++b) {
    const Block &B = S.blk[b];
    for (int k = B.s[2]; k <= B.e[2]; ++k)
      for (int j = B.s[1]; j <= B.e[1]; ++j)
        for (int i = B.s[0]; i <= B.e[0]; ++i)
          for (int q = 0; q < 3; ++q) S.Pc(0, b, S.iBX + q)[B.idx(k, j, i)] *= norm;
  }
  return 0;
}

// ---- fm_torus_cks (hydrodynamic: the reference input has enable_B = false)
// Pinned by tests/test_refinement.py: Fishbone-Moncrief's u^t u_phi is the
// same in every torus cell and equals the MKS torus's (init_torus above).
// fm_torus_cks/problem_generator.cpp:37-61: Boyer-Lindquist metric at (r, th).
static void bl_metric_rth(double r, double th, double a, double g[4][4]) {
  for (int row = 0; row < 4; ++row)
    for (int col = 0; col < 4; ++col) g[row][col] = 0.0;
  const double cth = std::cos(th);
  const double sth = std::sin(th);
  const double sth2 = sth * sth;
  const double r2 = r * r;
  const double a2 = a * a;
  const double mu = 1.0 + a2 * cth * cth / r2;
  const double delta = 1.0 - 2.0 / r + a2 / r2;
  const double inv_r_mu = 1.0 / (r * mu);
  g[0][0] = -1.0 + 2.0 * inv_r_mu;
  g[0][3] = -2.0 * a * sth2 * inv_r_mu;
  g[1][1] = mu / delta;
  g[2][2] = r2 * mu;
  g[3][0] = g[0][3];
  g[3][3] = r2 * sth2 * (1.0 + a2 / r2 + 2.0 * a2 * sth2 / (r * r2 * mu));
}

// Kerr-Schild radius clamped at 1 and polar angle of a CKS point (:82-90,
// :186-190).
static void ks_r_th(double x, double y, double z, double a2, double &r, double &th) {
  const double rad2 = x * x + y * y + z * z;
  r = std::fmax(std::sqrt((rad2 - a2 + std::sqrt((rad2 - a2) * (rad2 - a2) + 4.0 * a2 * z * z)) / 2.0), 1.0);
  th = (std::fabs(z / r) < 1.0) ? std::acos(z / r) : ((z > 0) ? 0.0 : M_PI);
}

// :76-141: BL four-velocity -> KS (radial terms) -> CKS Cartesian Jacobian.
static void bl_to_cks_ucon(double x, double y, double z, double a, const double ubl_in[4], double uc[4]) {
  const double a2 = a * a;
  double r, th;
  ks_r_th(x, y, z, a2, r, th);
  double gbl[4][4];
  bl_metric_rth(r, th, a, gbl);
  double ubl[4];
  for (int row = 0; row < 4; ++row) ubl[row] = ubl_in[row];
  ubl[0] = solve_ut(gbl, ubl[1], ubl[2], ubl[3]);
  const double delta = r * r - 2.0 * r + a2;
  double uks[4];
  uks[0] = ubl[0] + (2.0 * r / delta) * ubl[1];
  uks[1] = ubl[1];
  uks[2] = ubl[2];
  uks[3] = ubl[3] + (a / delta) * ubl[1];
  const double inv_r2pa2 = 1.0 / (r * r + a2);
  const double rho2 = x * x + y * y;
  const double sq = (rho2 > 1e-30) ? std::sqrt((r * r + a2) / rho2) : 0.0;
  double tr[4][4];
  for (int row = 0; row < 4; ++row)
    for (int col = 0; col < 4; ++col) tr[row][col] = 0.0;
  tr[0][0] = 1.0;
  tr[1][1] = (r * x + a * y) * inv_r2pa2;
  tr[1][2] = x * z / r * sq;
  tr[1][3] = (g_plant == PLANT_CKS_JACOBIAN) ? y : -y;
  tr[2][1] = (r * y - a * x) * inv_r2pa2;
  tr[2][2] = y * z / r * sq;
  tr[2][3] = x;
  tr[3][1] = z / r;
  tr[3][2] = -r * std::sqrt(rho2 * inv_r2pa2);
  for (int row = 0; row < 4; ++row) {
    uc[row] = 0.0;
    for (int col = 0; col < 4; ++col) uc[row] += tr[row][col] * uks[col];
  }
}

// ProblemGenerator (:143-334) over entire blocks, then
// MeshPostInitialization (:336-474): rho and u divided by the maximum
// density, then the radius-dependent floors (excision values inside
// r_excise, from the unclamped Boyer-Lindquist radius).  Entropy keeps its
// pre-normalisation value, as in the reference.
static int init_torus_cks(Sim &S) {
  const oracle_cfg &c = S.cfg;
  if (c.metric != MET_CKS) return -1;
  const double gam = c.gamma, a = c.a, a2 = a * a;
  const double rin = c.rin, kappa = c.kappa, pert = c.perturbation;
  const double l = lfish_calc(c.rmax, a);
  const double thin = M_PI_2;
  const double sthin = std::sin(thin), cthin = std::cos(thin);
  const double DDin = rin * rin - 2. * rin + a * a;
  const double AAin = (rin * rin + a * a) * (rin * rin + a * a) - DDin * a * a * sthin * sthin;
  const double SSin = rin * rin + a * a * cthin * cthin;
  int G[3], NX[3];
  for (int d = 0; d < 3; ++d) {
    G[d] = S.blk[0].ng[d];
    NX[d] = c.nx[d] + 2 * G[d];
  }
  for (int b = 0; b < S.nblk; ++b) {
    const Block &B = S.blk[b];
#pragma omp parallel for schedule(dynamic, 1)
    for (int k = 0; k < B.n[2]; ++k)
      for (int j = 0; j < B.n[1]; ++j)
        for (int i = 0; i < B.n[0]; ++i) {
          const int64_t id = B.idx(k, j, i);
          const int64_t gi = B.gofs[0] + i, gj = B.gofs[1] + j, gk = B.gofs[2] + k;
          const double rnd = keyed_uniform(c.seed, (uint64_t)((gk * NX[1] + gj) * NX[0] + gi));
          const double x = Xc(B, 0, i), y = Xc(B, 1, j), z = Xc(B, 2, k);
          const double xc[4] = {0.0, x, y, z};
          double r, th;
          ks_r_th(x, y, z, a2, r, th);
          const double sth = std::sin(th), cth = std::cos(th);
          double gcov[4][4], gcon[4][4], gbl[4][4];
          code_metric(MET_CKS, xc, gcov, c.h, a);
          bl_metric_rth(r, th, a, gbl);
          invert4(gcov, gcon);
          const double alpha = 1.0 / std::sqrt(-gcon[0][0]);
          const double beta1 = gcon[0][1] * alpha * alpha;
          const double beta2 = gcon[0][2] * alpha * alpha;
          const double beta3 = gcon[0][3] * alpha * alpha;
          double rho = 0.0, eint = 0.0, ent = 0.0, w1 = 0.0, w2 = 0.0, w3 = 0.0;
          const double DD = r * r - 2.0 * r + a2;
          const double AA = (r * r + a2) * (r * r + a2) - DD * a2 * sth * sth;
          const double SS = r * r + a2 * cth * cth;
          double lnh = 1.0;
          if (r >= rin) {
            lnh = 0.5 * std::log((1. + std::sqrt(1. + 4. * (l * l * SS * SS) * DD /
                                                         (AA * sth * AA * sth))) /
                                 (SS * DD / AA)) -
                  0.5 * std::sqrt(1. + 4. * (l * l * SS * SS) * DD / (AA * AA * sth * sth)) -
                  2. * a * r * l / AA -
                  (0.5 * std::log((1. + std::sqrt(1. + 4. * (l * l * SSin * SSin) * DDin /
                                                           (AAin * AAin * sthin * sthin))) /
                                  (SSin * DDin / AAin)) -
                   0.5 * std::sqrt(1. + 4. * (l * l * SSin * SSin) * DDin /
                                            (AAin * AAin * sthin * sthin)) -
                   2. * a * rin * l / AAin);
          }
          if (lnh >= 0.0 && r >= rin) {
            const double hm1 = std::exp(lnh) - 1.0;
            if (hm1 > 0.0) {
              rho = std::pow(hm1 * (gam - 1.0) / (kappa * gam), 1.0 / (gam - 1.0));
              eint = kappa * std::pow(rho, gam) / (gam - 1.0);
              eint *= (1.0 + pert * rnd);
              ent = (gam - 1) * eint * std::pow(rho, -gam);
              const double e2c = SS * SS * DD / (AA * AA * sth * sth);
              const double up1 = std::sqrt((-1.0 + std::sqrt(1.0 + 4.0 * l * l * e2c)) / 2.0);
              const double upbl = 2.0 * a * r * std::sqrt(1.0 + up1 * up1) / std::sqrt(AA * SS * DD) +
                                  std::sqrt(SS / AA) * up1 / sth;
              const double u0bl = solve_ut(gbl, 0.0, 0.0, upbl);
              const double ubl[4] = {u0bl, 0.0, 0.0, upbl};
              double uc[4];
              bl_to_cks_ucon(x, y, z, a, ubl, uc);
              w1 = uc[1] + beta1 * uc[0];
              w2 = uc[2] + beta2 * uc[0];
              w3 = uc[3] + beta3 * uc[0];
            }
          }
          S.Pc(0, b, 0)[id] = rho;
          S.Pc(0, b, 1)[id] = eint;
          S.Pc(0, b, 2)[id] = w1;
          S.Pc(0, b, 3)[id] = w2;
          S.Pc(0, b, 4)[id] = w3;
          S.Pc(0, b, S.iENT)[id] = ent;
          if (S.iKEL >= 0) S.Pc(0, b, S.iKEL)[id] = c.fel_0 * ent;
        }
  }
  double rhomax = 0.0;
  for (int b = 0; b < S.nblk; ++b)
    for (int64_t id = 0; id < S.blk[b].cells; ++id) rhomax = kmax(rhomax, S.Pc(0, b, 0)[id]);
  const double e_excise = c.pexcise / (gam - 1.0);
  for (int b = 0; b < S.nblk; ++b) {
    const Block &B = S.blk[b];
    for (int64_t id = 0; id < B.cells; ++id) {
      S.Pc(0, b, 0)[id] /= rhomax;
      S.Pc(0, b, 1)[id] /= rhomax;
    }
    for (int k = 0; k < B.n[2]; ++k)
      for (int j = 0; j < B.n[1]; ++j)
        for (int i = 0; i < B.n[0]; ++i) {
          const double xc = Xc(B, 0, i), yc = Xc(B, 1, j), zc = Xc(B, 2, k);
          const double rad2 = xc * xc + yc * yc + zc * zc;
          const double disc = (rad2 - a2) * (rad2 - a2) + 4.0 * a2 * zc * zc;
          const double r_bl = std::sqrt(0.5 * (rad2 - a2 + std::sqrt(disc)));
          double rho_floor, eng_floor;
          if (r_bl < c.r_excise) {
            rho_floor = c.dexcise;
            eng_floor = e_excise;
          } else {
            rho_floor = c.density_floor * std::pow(std::fmax(r_bl, 1.0), c.density_floor_pow);
            eng_floor = c.energy_floor * std::pow(std::fmax(r_bl, 1.0), c.energy_floor_pow);
          }
          const int64_t id = B.idx(k, j, i);
          if (S.Pc(0, b, 0)[id] < rho_floor) S.Pc(0, b, 0)[id] = rho_floor;
          if (S.Pc(0, b, 1)[id] < eng_floor) S.Pc(0, b, 1)[id] = eng_floor;
        }
  }
  if (!S.p.enable_B) return 0;
  // Magnetic field (:476-738).  Vector potential at cell centres over
  // entire blocks from the normalised, floored density (:519-545).
  for (int b = 0; b < S.nblk; ++b) {
    const Block &B = S.blk[b];
    std::vector<double> A1((size_t)B.cells), A2((size_t)B.cells);
    for (int k = 0; k < B.n[2]; ++k)
      for (int j = 0; j < B.n[1]; ++j)
        for (int i = 0; i < B.n[0]; ++i) {
          const double rho = S.Pc(0, b, 0)[B.idx(k, j, i)];
          double ascale = (rho > c.aphi_cut) ? (rho - c.aphi_cut) : 0.0;
          if (c.magnetic_mad && ascale > 0.0) {
            const double x = Xc(B, 0, i), y = Xc(B, 1, j), zc = Xc(B, 2, k);
            const double rad2 = x * x + y * y + zc * zc;
            const double r = std::fmax(
                std::sqrt((rad2 - a2 + std::sqrt((rad2 - a2) * (rad2 - a2) + 4.0 * a2 * zc * zc)) / 2.0), 1.0);
            const double cos_theta = std::fmin(std::fmax(zc / r, -1.0), 1.0);
            const double sin_theta = std::sqrt(1.0 - cos_theta * cos_theta);
            const double mad_scale = std::pow((r / c.rin) * sin_theta, 3.0) * std::exp(-r / 400.0);
            ascale *= mad_scale;
          }
          A1[(size_t)B.idx(k, j, i)] = -Xc(B, 1, j) * ascale;
          A2[(size_t)B.idx(k, j, i)] = Xc(B, 0, i) * ascale;
        }
    // B = curl A / sqrt|g| on the interior (:548-578)
    for (int k = B.s[2]; k <= B.e[2]; ++k)
      for (int j = B.s[1]; j <= B.e[1]; ++j)
        for (int i = B.s[0]; i <= B.e[0]; ++i) {
          const double xc[4] = {0.0, Xc(B, 0, i), Xc(B, 1, j), Xc(B, 2, k)};
          double gcov[4][4];
          code_metric(MET_CKS, xc, gcov, 0.0, a);
          const double sqrt_abs_gdet = std::sqrt(std::fabs(determinant4(gcov)));
          const double dx1 = B.dx[0], dx2 = B.dx[1], dx3 = B.dx[2];
          const double dA2_dx = (A2[(size_t)B.idx(k, j, i + 1)] - A2[(size_t)B.idx(k, j, i - 1)]) / (2.0 * dx1);
          const double dA1_dy = (A1[(size_t)B.idx(k, j + 1, i)] - A1[(size_t)B.idx(k, j - 1, i)]) / (2.0 * dx2);
          const double dA2_dz = (A2[(size_t)B.idx(k + 1, j, i)] - A2[(size_t)B.idx(k - 1, j, i)]) / (2.0 * dx3);
          const double dA1_dz = (A1[(size_t)B.idx(k + 1, j, i)] - A1[(size_t)B.idx(k - 1, j, i)]) / (2.0 * dx3);
          const int64_t id = B.idx(k, j, i);
          S.Pc(0, b, S.iBX)[id] = (g_plant == PLANT_CKS_CURL ? dA2_dz : -dA2_dz) / sqrt_abs_gdet;
          S.Pc(0, b, S.iBX + 1)[id] = dA1_dz / sqrt_abs_gdet;
          S.Pc(0, b, S.iBX + 2)[id] = (dA2_dx - dA1_dy) / sqrt_abs_gdet;
        }
  }
  // beta normalisation (:581-730): max b^2 over interiors, max u and rho
  // over entire blocks
  double bsq_max = 0.0, umax_n = 0.0, rhomax_n = 0.0;
  for (int b = 0; b < S.nblk; ++b) {
    const Block &B = S.blk[b];
    for (int k = B.s[2]; k <= B.e[2]; ++k)
      for (int j = B.s[1]; j <= B.e[1]; ++j)
        for (int i = B.s[0]; i <= B.e[0]; ++i) {
          const double xc[4] = {0.0, Xc(B, 0, i), Xc(B, 1, j), Xc(B, 2, k)};
          double gcov[4][4], gcon[4][4];
          code_metric(MET_CKS, xc, gcov, 0.0, a);
          invert4(gcov, gcon);
          real pc[NPRIM];
          S.load_prim(0, b, B.idx(k, j, i), pc);
          State st;
          calc_state(pc, gcov, gcon, st);
          bsq_max = kmax(bsq_max, st.bsq);
        }
    for (int64_t id = 0; id < B.cells; ++id) {
      umax_n = kmax(umax_n, S.Pc(0, b, 1)[id]);
      rhomax_n = kmax(rhomax_n, S.Pc(0, b, 0)[id]);
    }
  }
  const double beta_min = 2.0 * (gam - 1.0) * (umax_n / rhomax_n) / (bsq_max + 1.0e-30);
  const double magnetic_norm = std::sqrt(beta_min / c.beta);
  for (int b = 0; b < S.nblk; ++b) {
    const Block &B = S.blk[b];
    for (int k = B.s[2]; k <= B.e[2]; ++k)
      for (int j = B.s[1]; j <= B.e[1]; ++j)
        for (int i = B.s[0]; i <= B.e[0]; ++i)
          for (int q = 0; q < 3; ++q) S.Pc(0, b, S.iBX + q)[B.idx(k, j, i)] *= magnetic_norm;
  }
  return 0;
}

// ProblemGenerator + MeshPostInitialization, then the initial ghost exchange
// and physical boundary conditions of the "base" container.
int Sim::init_problem() {
  if (cfg.problem == 1)
    init_ot(*this);
  else if (cfg.problem == 2)
    init_torus(*this);
  else if (cfg.problem == 3) {
    if (init_torus_cks(*this)) return -1;
  } else
    return -1;
  exchange(0);
  apply_bc(0);
  return 0;
}

}  // namespace oracle
