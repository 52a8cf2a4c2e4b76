// Initial data for the "base" container (host, FP64, g++ -ffp-contract=off).
//
//  problem 1 - relativistic Orszag-Tang vortex on the unit box:
//    rho = 25/(36 pi), p = 5/(12 pi), v = 0.5(-sin 2 pi y, sin 2 pi x, 0),
//    B = (-sin 2 pi y, sin 4 pi x, 0)/sqrt(4 pi), Minkowski.
//  problem 2 - Fishbone-Moncrief torus in MKS with a SANE A_phi field and
//    beta normalisation (pangu/problem/fm_torus_mks/problem_generator.cpp).
//    The 4% pressure perturbation uses a counter-based generator keyed by
//    (seed, global cell index), so the data do not depend on thread order.
#include <cmath>
#include <cstring>
#include <functional>

#include "host.h"

namespace pangu {

HostBlock make_host_block(const pangu_block &b) {
  HostBlock h;
  for (int d = 0; d < 3; ++d) {
    h.nx[d] = b.nx[d];
    h.ng[d] = b.ng[d];
    h.n[d] = b.nx[d] + 2 * b.ng[d];
    h.s[d] = b.ng[d];
    h.e[d] = b.ng[d] + b.nx[d] - 1;
    h.lx[d] = b.lx[d];
    h.dx[d] = b.dx[d];
    h.xg[d] = b.xmin[d] - b.ng[d] * b.dx[d];
  }
  h.level = b.level;
  h.cells = (int64_t)h.n[0] * h.n[1] * h.n[2];
  return h;
}

static double keyed_uniform(uint64_t seed, uint64_t id) {
  uint64_t z = id + seed * 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  z ^= z >> 31;
  return -1.0 + 2.0 * ((double)(z >> 11) * (1.0 / 9007199254740992.0));
}

static inline double kmax(double a, double b) { return (a < b) ? b : a; }

// b^2 of a primitive state (state_calculation.h:19-69), host version used by
// the beta normalisation.
static double bsq_of(const double p[10], const geom::M4 g, const geom::M4 gi) {
  const double alpha = 1.0 / std::sqrt(-gi[0][0]);
  const double v[3] = {p[2], p[3], p[4]};
  double sq = 0.0;
  bool first = true;
  for (int a = 0; a < 3; ++a)
    for (int c = 0; c < 3; ++c) {
      const double t = v[a] * g[a + 1][c + 1] * v[c];
      sq = first ? t : sq + t;
      first = false;
    }
  const double lor = std::sqrt(1.0 + sq);
  double uc[4], ul[4], bc[4], bl[4];
  uc[0] = lor / alpha;
  for (int a = 1; a < 4; ++a) uc[a] = v[a - 1] - uc[0] * gi[0][a] * alpha * alpha;
  for (int r = 0; r < 4; ++r) ul[r] = g[r][0] * uc[0] + g[r][1] * uc[1] + g[r][2] * uc[2] + g[r][3] * uc[3];
  bc[0] = p[5] * ul[1] + p[6] * ul[2] + p[7] * ul[3];
  for (int a = 1; a < 4; ++a) bc[a] = (p[4 + a] + bc[0] * uc[a]) / uc[0];
  for (int r = 0; r < 4; ++r) bl[r] = g[r][0] * bc[0] + g[r][1] * bc[1] + g[r][2] * bc[2] + g[r][3] * bc[3];
  return bc[0] * bl[0] + bc[1] * bl[1] + bc[2] * bl[2] + bc[3] * bl[3];
}

// fm_torus_mks/problem_generator.cpp:22-35 (specific angular momentum).
static double fish_l(double r, double a) {
  const double sr = std::sqrt(r);
  const double q = std::pow(a, 2.) - 2. * a * sr + std::pow(r, 2.);
  const double s1 = std::sqrt(2. * a * sr + (-3. + r) * r);
  const double t1 = (-2. * a * r * q) / s1;
  const double t2 = ((a + (-2. + r) * sr) * (std::pow(r, 3.) + std::pow(a, 2.) * (2. + r))) /
                    std::sqrt(1 + (2. * a) / std::pow(r, 1.5) - 3. / r);
  return (q * (t1 + t2)) / (std::pow(r, 3.) * s1 * (std::pow(a, 2.) + (-2. + r) * r));
}

// ln(h) of the FM torus up to the inner-edge constant (:208-222).
static double fish_lnh_part(double l, double r, double a, double SS, double DD, double AA, double sth) {
  return 0.5 * std::log((1. + std::sqrt(1. + 4. * (l * l * SS * SS) * DD / (AA * sth * AA * sth))) /
                        (SS * DD / AA)) -
         0.5 * std::sqrt(1. + 4. * (l * l * SS * SS) * DD / (AA * AA * sth * sth)) -
         2. * a * r * l / AA;
}

static void ot_cell(const pangu_config &cfg, double x, double y, double p[10]) {
  const double gam = cfg.gamma;
  const double pi = M_PI;
  const double rho = 25.0 / (36.0 * pi);
  const double u = (5.0 / (12.0 * pi)) / (gam - 1.0);
  const double vx = -0.5 * std::sin(2.0 * pi * y);
  const double vy = 0.5 * std::sin(2.0 * pi * x);
  const double W = 1.0 / std::sqrt(1.0 - (vx * vx + vy * vy));
  const double b0 = 1.0 / std::sqrt(4.0 * pi);
  p[0] = rho;
  p[1] = u;
  p[2] = W * vx;
  p[3] = W * vy;
  p[4] = 0.0;
  p[5] = -b0 * std::sin(2.0 * pi * y);
  p[6] = b0 * std::sin(4.0 * pi * x);
  p[7] = 0.0;
  p[8] = (gam - 1.0) * u * std::pow(rho, -gam);
}

// fm_torus_cks/problem_generator.cpp:37-61: Boyer-Lindquist metric at (r, th).
static void bl_metric_rth(double r, double th, double a, geom::M4 g) {
  for (int row = 0; row < 4; ++row)
    for (int col = 0; col < 4; ++col) g[row][col] = 0.0;
  const double cth = std::cos(th), sth = std::sin(th), sth2 = sth * sth;
  const double r2 = r * r, a2 = a * a;
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

// u^t of a BL four-velocity with u^r = u^theta = 0 (:63-74).
static double bl_ut(const geom::M4 g, double u3) {
  const double AA = g[0][0];
  const double BB = 2.0 * (g[0][1] * 0.0 + g[0][2] * 0.0 + g[0][3] * u3);
  const double CC = 1.0 + g[1][1] * 0.0 * 0.0 + g[2][2] * 0.0 * 0.0 + g[3][3] * u3 * u3 +
                    2.0 * (g[1][2] * 0.0 * 0.0 + g[1][3] * 0.0 * u3 + g[2][3] * 0.0 * u3);
  return (-BB - std::sqrt(BB * BB - 4.0 * AA * CC)) / (2.0 * AA);
}

// Kerr-Schild radius (clamped at 1) and polar angle of a CKS point (:82-90).
static void ks_r_th(double x, double y, double z, double a2, double &r, double &th) {
  const double rad2 = x * x + y * y + z * z;
  r = std::fmax(std::sqrt((rad2 - a2 + std::sqrt((rad2 - a2) * (rad2 - a2) + 4.0 * a2 * z * z)) / 2.0), 1.0);
  th = (std::fabs(z / r) < 1.0) ? std::acos(z / r) : ((z > 0) ? 0.0 : M_PI);
}

// FM torus on Cartesian Kerr-Schild coordinates without a magnetic field
// (fm_torus_cks/problem_generator.cpp:143-474): the BL torus state, its
// four-velocity taken to KS and then to the Cartesian basis (:76-141), rho
// and u divided by the maximum density, then radius-dependent floors
// (excision values inside r_excise).  Entropy keeps its pre-normalisation
// value, as the reference leaves it.
static int torus_cks(const pangu_config &cfg, const ProblemParams &pp,
                     const std::vector<HostBlock> &blocks, int iBX, int iENT, int iKEL,
                     const std::function<double *(size_t, int)> &P) {
  const double gam = cfg.gamma, a = cfg.a, a2 = a * a;
  const double rin = pp.rin, kap = pp.kappa;
  const double l = fish_l(pp.rmax, a);
  const double sthin = std::sin(M_PI_2), cthin = std::cos(M_PI_2);
  const double DDin = rin * rin - 2. * rin + a * a;
  const double AAin = (rin * rin + a * a) * (rin * rin + a * a) - DDin * a * a * sthin * sthin;
  const double SSin = rin * rin + a * a * cthin * cthin;
  const double lnh_in = fish_lnh_part(l, rin, a, SSin, DDin, AAin, sthin);
  int NX[3];
  for (int d = 0; d < 3; ++d) NX[d] = pp.mesh_nx[d] + 2 * blocks[0].ng[d];
  for (size_t b = 0; b < blocks.size(); ++b) {
    const HostBlock &B = blocks[b];
#pragma omp parallel for schedule(dynamic, 1)
    for (int k = 0; k < B.n[2]; ++k)
      for (int j = 0; j < B.n[1]; ++j)
        for (int i = 0; i < B.n[0]; ++i) {
          const int64_t gi = (int64_t)B.lx[0] * B.nx[0] + i, gj = (int64_t)B.lx[1] * B.nx[1] + j,
                        gk = (int64_t)B.lx[2] * B.nx[2] + k;
          const double rnd = keyed_uniform(pp.seed, (uint64_t)((gk * NX[1] + gj) * NX[0] + gi));
          const double x = B.xc(0, i), y = B.xc(1, j), z = B.xc(2, k);
          const double xc[4] = {0.0, x, y, z};
          double r, th;
          ks_r_th(x, y, z, a2, r, th);
          const double sth = std::sin(th), cth = std::cos(th);
          geom::M4 g, gi_, gbl;
          geom::code_metric(PANGU_METRIC_CKS, xc, g, cfg.h, a);
          bl_metric_rth(r, th, a, gbl);
          geom::inv4(g, gi_);
          const double alpha = 1.0 / std::sqrt(-gi_[0][0]);
          double rho = 0.0, eint = 0.0, ent = 0.0, w[3] = {0.0, 0.0, 0.0};
          const double DD = r * r - 2.0 * r + a2;
          const double AA = (r * r + a2) * (r * r + a2) - DD * a2 * sth * sth;
          const double SS = r * r + a2 * cth * cth;
          double lnh = 1.0;
          if (r >= rin) lnh = fish_lnh_part(l, r, a, SS, DD, AA, sth) - lnh_in;
          if (lnh >= 0.0 && r >= rin) {
            const double hm1 = std::exp(lnh) - 1.0;
            if (hm1 > 0.0) {
              rho = std::pow(hm1 * (gam - 1.0) / (kap * gam), 1.0 / (gam - 1.0));
              eint = kap * std::pow(rho, gam) / (gam - 1.0);
              eint *= (1.0 + pp.perturbation * rnd);
              ent = (gam - 1) * eint * std::pow(rho, -gam);
              const double e2c = SS * SS * DD / (AA * AA * sth * sth);
              const double up1 = std::sqrt((-1.0 + std::sqrt(1.0 + 4.0 * l * l * e2c)) / 2.0);
              const double upbl = 2.0 * a * r * std::sqrt(1.0 + up1 * up1) / std::sqrt(AA * SS * DD) +
                                  std::sqrt(SS / AA) * up1 / sth;
              // BL -> KS: u^t += 2r/Delta u^r, u^phi += a/Delta u^r with u^r = 0
              const double u0 = bl_ut(gbl, upbl);
              const double delta = r * r - 2.0 * r + a2;
              const double ks[4] = {u0 + (2.0 * r / delta) * 0.0, 0.0, 0.0, upbl + (a / delta) * 0.0};
              // KS -> Cartesian KS Jacobian
              const double inv_r2pa2 = 1.0 / (r * r + a2);
              const double rho2 = x * x + y * y;
              const double sq = (rho2 > 1e-30) ? std::sqrt((r * r + a2) / rho2) : 0.0;
              geom::M4 tr;
              for (int rr = 0; rr < 4; ++rr)
                for (int cc = 0; cc < 4; ++cc) tr[rr][cc] = 0.0;
              tr[0][0] = 1.0;
              tr[1][1] = (r * x + a * y) * inv_r2pa2;
              tr[1][2] = x * z / r * sq;
              tr[1][3] = -y;
              tr[2][1] = (r * y - a * x) * inv_r2pa2;
              tr[2][2] = y * z / r * sq;
              tr[2][3] = x;
              tr[3][1] = z / r;
              tr[3][2] = -r * std::sqrt(rho2 * inv_r2pa2);
              double uc[4];
              for (int rr = 0; rr < 4; ++rr) {
                uc[rr] = 0.0;
                for (int cc = 0; cc < 4; ++cc) uc[rr] += tr[rr][cc] * ks[cc];
              }
              for (int q = 0; q < 3; ++q) w[q] = uc[q + 1] + (gi_[0][q + 1] * alpha * alpha) * uc[0];
            }
          }
          const int64_t id = B.id(k, j, i);
          P(b, 0)[id] = rho;
          P(b, 1)[id] = eint;
          P(b, 2)[id] = w[0];
          P(b, 3)[id] = w[1];
          P(b, 4)[id] = w[2];
          P(b, iENT)[id] = ent;
          if (iKEL >= 0) P(b, iKEL)[id] = pp.fel_0 * ent;
        }
  }
  double rhomax = 0.0;
  for (size_t b = 0; b < blocks.size(); ++b)
    for (int64_t id = 0; id < blocks[b].cells; ++id) rhomax = kmax(rhomax, P(b, 0)[id]);
  const double e_excise = cfg.pexcise / (gam - 1.0);
  for (size_t b = 0; b < blocks.size(); ++b) {
    const HostBlock &B = blocks[b];
    for (int64_t id = 0; id < B.cells; ++id) {
      P(b, 0)[id] /= rhomax;
      P(b, 1)[id] /= rhomax;
    }
    for (int k = 0; k < B.n[2]; ++k)
      for (int j = 0; j < B.n[1]; ++j)
        for (int i = 0; i < B.n[0]; ++i) {
          const double xc = B.xc(0, i), yc = B.xc(1, j), zc = B.xc(2, k);
          const double rad2 = xc * xc + yc * yc + zc * zc;
          const double disc = (rad2 - a2) * (rad2 - a2) + 4.0 * a2 * zc * zc;
          const double r_bl = std::sqrt(0.5 * (rad2 - a2 + std::sqrt(disc)));
          double rf, ef;
          if (r_bl < cfg.r_excise) {
            rf = cfg.dexcise;
            ef = e_excise;
          } else {
            rf = cfg.density_floor * std::pow(std::fmax(r_bl, 1.0), cfg.density_floor_pow);
            ef = cfg.energy_floor * std::pow(std::fmax(r_bl, 1.0), cfg.energy_floor_pow);
          }
          const int64_t id = B.id(k, j, i);
          if (P(b, 0)[id] < rf) P(b, 0)[id] = rf;
          if (P(b, 1)[id] < ef) P(b, 1)[id] = ef;
        }
  }
  if (iBX < 0) return 0;
  // Field (:476-738): A = (-y, x, 0) max(rho - cut, 0) at cell centres (MAD:
  // times ((r / rin) sin(theta))^3 exp(-r / 400)), B = curl A / sqrt|g| by
  // central differences over the interior, then scaled so that
  // 2 (gamma - 1) (u_max / rho_max) / max b^2 = beta.
  for (size_t b = 0; b < blocks.size(); ++b) {
    const HostBlock &B = blocks[b];
    std::vector<double> A1((size_t)B.cells), A2((size_t)B.cells);
    for (int k = 0; k < B.n[2]; ++k)
      for (int j = 0; j < B.n[1]; ++j)
        for (int i = 0; i < B.n[0]; ++i) {
          const int64_t id = B.id(k, j, i);
          const double rho = P(b, 0)[id];
          double ascale = (rho > pp.aphi_cut) ? (rho - pp.aphi_cut) : 0.0;
          if (pp.mad && ascale > 0.0) {
            const double x = B.xc(0, i), y = B.xc(1, j), zc = B.xc(2, k);
            const double rad2 = x * x + y * y + zc * zc;
            const double r = std::fmax(
                std::sqrt((rad2 - a2 + std::sqrt((rad2 - a2) * (rad2 - a2) + 4.0 * a2 * zc * zc)) / 2.0), 1.0);
            const double cth = std::fmin(std::fmax(zc / r, -1.0), 1.0);
            const double sth = std::sqrt(1.0 - cth * cth);
            ascale *= std::pow((r / rin) * sth, 3.0) * std::exp(-r / 400.0);
          }
          A1[(size_t)id] = -B.xc(1, j) * ascale;
          A2[(size_t)id] = B.xc(0, i) * ascale;
        }
#pragma omp parallel for schedule(dynamic, 1)
    for (int k = B.s[2]; k <= B.e[2]; ++k)
      for (int j = B.s[1]; j <= B.e[1]; ++j)
        for (int i = B.s[0]; i <= B.e[0]; ++i) {
          const double xc[4] = {0.0, B.xc(0, i), B.xc(1, j), B.xc(2, k)};
          geom::M4 g;
          geom::code_metric(PANGU_METRIC_CKS, xc, g, 0.0, a);
          const double sg = std::sqrt(std::fabs(geom::det4(g)));
          const int64_t id = B.id(k, j, i), sj = B.n[0], sk = (int64_t)B.n[0] * B.n[1];
          const double dA2_dx = (A2[(size_t)(id + 1)] - A2[(size_t)(id - 1)]) / (2.0 * B.dx[0]);
          const double dA1_dy = (A1[(size_t)(id + sj)] - A1[(size_t)(id - sj)]) / (2.0 * B.dx[1]);
          const double dA2_dz = (A2[(size_t)(id + sk)] - A2[(size_t)(id - sk)]) / (2.0 * B.dx[2]);
          const double dA1_dz = (A1[(size_t)(id + sk)] - A1[(size_t)(id - sk)]) / (2.0 * B.dx[2]);
          P(b, iBX)[id] = -dA2_dz / sg;
          P(b, iBX + 1)[id] = dA1_dz / sg;
          P(b, iBX + 2)[id] = (dA2_dx - dA1_dy) / sg;
        }
  }
  double bsqmax = 0.0, umax = 0.0, rmax_n = 0.0;
  for (size_t b = 0; b < blocks.size(); ++b) {
    const HostBlock &B = blocks[b];
    double bm = 0.0;
    // max of non-negative values: order-independent
#pragma omp parallel for schedule(dynamic, 1) reduction(max : bm)
    for (int k = B.s[2]; k <= B.e[2]; ++k)
      for (int j = B.s[1]; j <= B.e[1]; ++j)
        for (int i = B.s[0]; i <= B.e[0]; ++i) {
          const double xc[4] = {0.0, B.xc(0, i), B.xc(1, j), B.xc(2, k)};
          geom::M4 g, gi_;
          geom::code_metric(PANGU_METRIC_CKS, xc, g, 0.0, a);
          geom::inv4(g, gi_);
          const int64_t id = B.id(k, j, i);
          double p[10] = {P(b, 0)[id], P(b, 1)[id], P(b, 2)[id], P(b, 3)[id], P(b, 4)[id],
                          P(b, iBX)[id], P(b, iBX + 1)[id], P(b, iBX + 2)[id], 0.0, 0.0};
          bm = kmax(bm, bsq_of(p, g, gi_));
        }
    bsqmax = kmax(bsqmax, bm);
    for (int64_t id = 0; id < B.cells; ++id) {
      umax = kmax(umax, P(b, 1)[id]);
      rmax_n = kmax(rmax_n, P(b, 0)[id]);
    }
  }
  const double beta_min = 2.0 * (gam - 1.0) * (umax / rmax_n) / (bsqmax + 1.0e-30);
  const double norm = std::sqrt(beta_min / pp.beta);
  for (size_t b = 0; b < blocks.size(); ++b) {
    const HostBlock &B = blocks[b];
    for (int k = B.s[2]; k <= B.e[2]; ++k)
      for (int j = B.s[1]; j <= B.e[1]; ++j)
        for (int i = B.s[0]; i <= B.e[0]; ++i)
          for (int q = 0; q < 3; ++q) P(b, iBX + q)[B.id(k, j, i)] *= norm;
  }
  return 0;
}

int generate_problem(const pangu_config &cfg, const ProblemParams &pp,
                     const std::vector<HostBlock> &blocks, int np, int iBX, int iENT, int iKEL,
                     std::vector<double> &prim) {
  int64_t total = 0;
  std::vector<int64_t> off(blocks.size());
  for (size_t b = 0; b < blocks.size(); ++b) {
    off[b] = total;
    total += blocks[b].cells;
  }
  prim.assign((size_t)(total * np), 0.0);
  auto P = [&](size_t b, int q) { return prim.data() + off[b] * np + (int64_t)q * blocks[b].cells; };

  if (pp.problem == 1) {
    for (size_t b = 0; b < blocks.size(); ++b) {
      const HostBlock &B = blocks[b];
      for (int k = 0; k < B.n[2]; ++k)
        for (int j = 0; j < B.n[1]; ++j)
          for (int i = 0; i < B.n[0]; ++i) {
            double p[10];
            ot_cell(cfg, B.xc(0, i), B.xc(1, j), p);
            const int64_t id = B.id(k, j, i);
            for (int q = 0; q < 5; ++q) P(b, q)[id] = p[q];
            if (iBX >= 0)
              for (int q = 0; q < 3; ++q) P(b, iBX + q)[id] = p[5 + q];
            P(b, iENT)[id] = p[8];
            if (iKEL >= 0) P(b, iKEL)[id] = pp.fel_0 * p[8];
          }
    }
    return 0;
  }
  if (pp.problem == 3) {
    if (cfg.metric != PANGU_METRIC_CKS) return -1;
    return torus_cks(cfg, pp, blocks, iBX, iENT, iKEL, P);
  }
  if (pp.problem != 2 || cfg.metric != PANGU_METRIC_MKS) return -1;

  const double gam = cfg.gamma, a = cfg.a, h = cfg.h, a2 = a * a;
  const double rin = pp.rin, kap = pp.kappa;
  const double l = fish_l(pp.rmax, a);
  const double sthin = std::sin(M_PI_2), cthin = std::cos(M_PI_2);
  const double DDin = rin * rin - 2. * rin + a * a;
  const double AAin = (rin * rin + a * a) * (rin * rin + a * a) - DDin * a * a * sthin * sthin;
  const double SSin = rin * rin + a * a * cthin * cthin;
  const double lnh_in = fish_lnh_part(l, rin, a, SSin, DDin, AAin, sthin);
  int G[3], NX[3];
  for (int d = 0; d < 3; ++d) {
    G[d] = blocks[0].ng[d];
    NX[d] = pp.mesh_nx[d] + 2 * G[d];
  }
  for (size_t b = 0; b < blocks.size(); ++b) {
    const HostBlock &B = blocks[b];
#pragma omp parallel for schedule(dynamic, 1)
    for (int k = 0; k < B.n[2]; ++k)
      for (int j = 0; j < B.n[1]; ++j)
        for (int i = 0; i < B.n[0]; ++i) {
          const int64_t gi = (int64_t)B.lx[0] * B.nx[0] + i, gj = (int64_t)B.lx[1] * B.nx[1] + j,
                        gk = (int64_t)B.lx[2] * B.nx[2] + k;
          const double rnd = keyed_uniform(pp.seed, (uint64_t)((gk * NX[1] + gj) * NX[0] + gi));
          const double xc[4] = {0.0, B.xc(0, i), B.xc(1, j), B.xc(2, k)};
          const double r = std::exp(xc[1]);
          const double th = M_PI_2 * (xc[2] + 1.0) + 0.5 * h * std::sin(M_PI * (xc[2] + 1.0));
          const double sth = std::sin(th), cth = std::cos(th);
          geom::M4 g, gbl, gi_;
          geom::code_metric(PANGU_METRIC_MKS, xc, g, h, a);
          geom::code_metric(PANGU_METRIC_BL, xc, gbl, h, a);
          geom::inv4(g, gi_);
          const double alpha = 1.0 / std::sqrt(-gi_[0][0]);
          double rho = 0.0, eint = 0.0, ent = 0.0, w[3] = {0.0, 0.0, 0.0};
          const double DD = r * r - 2.0 * r + a2;
          const double AA = (r * r + a2) * (r * r + a2) - DD * a2 * sth * sth;
          const double SS = r * r + a2 * cth * cth;
          double lnh = 1.0;
          if (r >= rin) lnh = fish_lnh_part(l, r, a, SS, DD, AA, sth) - lnh_in;
          if (lnh >= 0.0 && r >= rin) {
            const double hm1 = std::exp(lnh) - 1.0;
            if (hm1 > 0.0) {
              rho = std::pow(hm1 * (gam - 1.0) / (kap * gam), 1.0 / (gam - 1.0));
              eint = kap * std::pow(rho, gam) / (gam - 1.0);
              eint *= (1.0 + pp.perturbation * rnd);
              ent = (gam - 1) * eint * std::pow(rho, -gam);
              const double e2c = SS * SS * DD / (AA * AA * sth * sth);
              const double up1 = std::sqrt((-1.0 + std::sqrt(1.0 + 4.0 * l * l * e2c)) / 2.0);
              const double upbl = 2.0 * a * r * std::sqrt(1.0 + up1 * up1) / std::sqrt(AA * SS * DD) +
                                  std::sqrt(SS / AA) * up1 / sth;
              // u^t from g_tt u^t^2 + 2 g_t3 u^t u^3 + (1 + g_33 u^3^2) = 0 (:37-48)
              const double AAq = gbl[0][0];
              const double BBq = 2.0 * (gbl[0][1] * 0.0 + gbl[0][2] * 0.0 + gbl[0][3] * upbl);
              const double CCq = 1.0 + gbl[1][1] * 0.0 * 0.0 + gbl[2][2] * 0.0 * 0.0 +
                                 gbl[3][3] * upbl * upbl +
                                 2.0 * (gbl[1][2] * 0.0 * 0.0 + gbl[1][3] * 0.0 * upbl +
                                        gbl[2][3] * 0.0 * upbl);
              const double u0bl = (-BBq - std::sqrt(BBq * BBq - 4.0 * AAq * CCq)) / (2.0 * AAq);
              // BL -> KS (:63-76): with u^r = u^theta = 0 the transformation
              // leaves (u^t, 0, 0, u^phi) unchanged; then into the code basis
              // with the inverse FD Jacobian of the coordinate map (:78-100).
              const double ks[4] = {u0bl, 0.0, 0.0, upbl};
              geom::M4 J, Ji;
              for (int c = 0; c < 4; ++c) {
                double xp[4], xm[4];
                std::memcpy(xp, xc, sizeof(xp));
                std::memcpy(xm, xc, sizeof(xm));
                xp[c] += 1e-5;
                xm[c] -= 1e-5;
                const double yp[4] = {xp[0], std::exp(xp[1]),
                                      M_PI_2 * (xp[2] + 1.0) + 0.5 * h * std::sin(M_PI * (xp[2] + 1.0)),
                                      xp[3]};
                const double ym[4] = {xm[0], std::exp(xm[1]),
                                      M_PI_2 * (xm[2] + 1.0) + 0.5 * h * std::sin(M_PI * (xm[2] + 1.0)),
                                      xm[3]};
                for (int rr = 0; rr < 4; ++rr) J[rr][c] = (yp[rr] - ym[rr]) / (xp[c] - xm[c]);
              }
              geom::inv4(J, Ji);
              double uc[4];
              for (int rr = 0; rr < 4; ++rr)
                uc[rr] = Ji[rr][0] * ks[0] + Ji[rr][1] * ks[1] + Ji[rr][2] * ks[2] + Ji[rr][3] * ks[3];
              for (int q = 0; q < 3; ++q) w[q] = uc[q + 1] + (gi_[0][q + 1] * alpha * alpha) * uc[0];
            }
          }
          const int64_t id = B.id(k, j, i);
          P(b, 0)[id] = rho;
          P(b, 1)[id] = eint;
          P(b, 2)[id] = w[0];
          P(b, 3)[id] = w[1];
          P(b, 4)[id] = w[2];
          P(b, iENT)[id] = ent;
          if (iKEL >= 0) P(b, iKEL)[id] = pp.fel_0 * ent;
        }
  }
  // maxima over entire blocks, then rho, u normalisation (:315-390)
  double rhomax = 0.0, umax = 0.0;
  for (size_t b = 0; b < blocks.size(); ++b) {
    double bm = 0.0, um = 0.0;
    for (int64_t id = 0; id < blocks[b].cells; ++id) {
      bm = kmax(bm, P(b, 0)[id]);
      um = kmax(um, P(b, 1)[id]);
    }
    rhomax = kmax(rhomax, bm);
    umax = kmax(umax, um);
  }
  // No cell inside the torus (an unbound torus, r_in inside the marginally
  // bound orbit for the spin, or one the mesh does not resolve): the
  // reference's division by the zero maximum fills the mesh with NaN; the
  // mesh is left as the zero state instead, which the first FixPrimitive
  // floors to the atmosphere, and no field is made.
  if (!(rhomax > 0.0)) return 0;
  for (size_t b = 0; b < blocks.size(); ++b)
    for (int64_t id = 0; id < blocks[b].cells; ++id) {
      P(b, 0)[id] /= rhomax;
      P(b, 1)[id] /= rhomax;
    }
  if (iBX < 0) return 0;
  // SANE A_phi at corners and B = curl A on the interior (:394-458)
  for (size_t b = 0; b < blocks.size(); ++b) {
    const HostBlock &B = blocks[b];
    std::vector<double> A((size_t)B.cells, 0.0);
    const double *rho = P(b, 0);
    for (int k = 0; k < B.n[2]; ++k)
      for (int j = 1; j < B.n[1]; ++j)
        for (int i = 1; i < B.n[0]; ++i) {
          const double avg =
              0.25 * (rho[B.id(k, j, i)] + rho[B.id(k, j, i - 1)] + rho[B.id(k, j - 1, i)] + rho[B.id(k, j - 1, i - 1)]);
          const double ex = avg - pp.aphi_cut;
          A[B.id(k, j, i)] = (ex > 0.0) ? ex : 0.0;
        }
#pragma omp parallel for schedule(dynamic, 1)
    for (int k = B.s[2]; k <= B.e[2]; ++k)
      for (int j = B.s[1]; j <= B.e[1]; ++j)
        for (int i = B.s[0]; i <= B.e[0]; ++i) {
          const double xc[4] = {0.0, B.xc(0, i), B.xc(1, j), B.xc(2, k)};
          geom::M4 g;
          geom::code_metric(PANGU_METRIC_MKS, xc, g, h, a);
          const double sg = std::sqrt(std::fabs(geom::det4(g)));
          const double a00 = A[B.id(k, j - 1, i - 1)], a01 = A[B.id(k, j, i - 1)];
          const double a10 = A[B.id(k, j - 1, i)], a11 = A[B.id(k, j, i)];
          P(b, iBX)[B.id(k, j, i)] = -(a00 - a01 + a10 - a11) / (2.0 * B.dx[1] * sg);
          P(b, iBX + 1)[B.id(k, j, i)] = (a00 + a01 - a10 - a11) / (2.0 * B.dx[0] * sg);
        }
  }
  // beta normalisation (:464-551)
  double bsqmax = 0.0;
  for (size_t b = 0; b < blocks.size(); ++b) {
    const HostBlock &B = blocks[b];
    double bm = 0.0;
    // max of non-negative values: order-independent, so a parallel reduction
    // gives the serial result
#pragma omp parallel for schedule(dynamic, 1) reduction(max : bm)
    for (int k = B.s[2]; k <= B.e[2]; ++k)
      for (int j = B.s[1]; j <= B.e[1]; ++j)
        for (int i = B.s[0]; i <= B.e[0]; ++i) {
          const double xc[4] = {0.0, B.xc(0, i), B.xc(1, j), B.xc(2, k)};
          geom::M4 g, gi_;
          geom::code_metric(PANGU_METRIC_MKS, xc, g, h, a);
          geom::inv4(g, gi_);
          const int64_t id = B.id(k, j, i);
          double p[10] = {P(b, 0)[id], P(b, 1)[id], P(b, 2)[id], P(b, 3)[id], P(b, 4)[id],
                          P(b, iBX)[id], P(b, iBX + 1)[id], P(b, iBX + 2)[id], 0.0, 0.0};
          bm = kmax(bm, bsq_of(p, g, gi_));
        }
    bsqmax = kmax(bsqmax, bm);
  }
  const double beta_min = 2.0 * (gam - 1.0) * (umax / rhomax) / (bsqmax + 1.0e-30);
  const double norm = std::sqrt(beta_min / pp.beta);
  for (size_t b = 0; b < blocks.size(); ++b) {
    const HostBlock &B = blocks[b];
    for (int k = B.s[2]; k <= B.e[2]; ++k)
      for (int j = B.s[1]; j <= B.e[1]; ++j)
        for (int i = B.s[0]; i <= B.e[0]; ++i)
          for (int q = 0; q < 3; ++q) P(b, iBX + q)[B.id(k, j, i)] *= norm;
  }
  return 0;
}

}  // namespace pangu
