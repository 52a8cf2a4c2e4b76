// Run diagnostics of the oracle state.  TEST INFRASTRUCTURE ONLY.
//
// The reference computes no diagnostics in its code; doc/grmhd_intro.md:556-574
// defines the usual ones, as integrals over a sphere of radius r:
//   Mdot  = -int rho u^r sqrt(-g) dtheta dphi
//   Edot  =  int T^r_t sqrt(-g) dtheta dphi
//   Jdot  = -int T^r_phi sqrt(-g) dtheta dphi
//   Phi_BH = 1/2 int |B^r| sqrt(-g) dtheta dphi
// In code coordinates the surface x1 = const carries the same integrals with
// u^1, T^1_0, T^1_3, B^1 and the code sqrt|g| over dx2 dx3 (a flux through a
// surface does not depend on the coordinates).  They are evaluated at the
// cell centres of one interior x1 index (the shell), plus
//   mass   = sum U_RHO dV,  energy = sum U_ENY dV   (interior cells)
//   divB   = max corner |div U_B| (the CT invariant, as in tests/helpers.py)
//   UBmax  = max |U_B| over the interior and the active dimensions.
// On a Cartesian Kerr-Schild mesh the cells Recovery excises
// (invertor.cc:83-91, r_BL < r_excise) hold no evolved state and their U can
// be non-finite: they are left out of mass, energy and UBmax, and corners
// touching one are left out of divB.
#include <cmath>

#include "oracle_sim.h"

namespace oracle {

void Sim::diagnostics(int c, int shell, double out[ODIAG_N], double mag[ODIAG_N]) const {
  for (int q = 0; q < ODIAG_N; ++q) out[q] = mag[q] = 0.0;
  // mag[q]: sum of |contribution| of each summed quantity (its rounding scale)
  auto add = [&](int q, double v) {
    out[q] += v;
    mag[q] += std::fabs(v);
  };
  const double a2 = p.a * p.a;
  auto excised = [&](const Block &B, int k, int j, int i) {
    if (p.metric != MET_CKS) return false;
    const real x = Xc(B, 0, i), y = Xc(B, 1, j), z = Xc(B, 2, k);
    const real rad2 = x * x + y * y + z * z;
    const real sa = (rad2 - a2) * (rad2 - a2) + 4.0 * a2 * z * z;
    return std::sqrt(0.5 * (rad2 - a2 + std::sqrt(sa))) < p.r_excise;
  };
  for (int b = 0; b < nblk; ++b) {
    const Block &B = blk[b];
    const double dA = B.dx[1] * B.dx[2];
    const double dV = B.dx[0] * B.dx[1] * B.dx[2];
    const double *u_rho = U[c].data() + B.off * nc;
    const double *u_eny = u_rho + B.cells;
    for (int k = B.s[2]; k <= B.e[2]; ++k)
      for (int j = B.s[1]; j <= B.e[1]; ++j)
        for (int i = B.s[0]; i <= B.e[0]; ++i) {
          const int64_t id = B.idx(k, j, i);
          if (excised(B, k, j, i)) continue;
          add(ODIAG_MASS, u_rho[id] * dV);
          add(ODIAG_ENERGY, u_eny[id] * dV);
        }
    const int i_sh = shell - B.lx[0] * B.nx[0];
    if (i_sh >= 0 && i_sh < B.nx[0]) {
      const int i = B.s[0] + i_sh;
      for (int k = B.s[2]; k <= B.e[2]; ++k)
        for (int j = B.s[1]; j <= B.e[1]; ++j) {
          const int64_t id = B.idx(k, j, i);
          real gcov[4][4], gcon[4][4], gdet;
          metric(b, 0, k, j, i, gcov, gcon, &gdet);
          real pc[NPRIM];
          load_prim(c, b, id, pc);
          State s;
          calc_state(pc, gcov, gcon, s);
          real t[4];
          tmunu_row(p.gamma, pc, gcov, gcon, 1, t);
          const real sg = std::sqrt(std::fabs(gdet));
          add(ODIAG_MDOT, -(pc[RHO] * s.ucon[1] * sg) * dA);
          add(ODIAG_EDOT, t[0] * sg * dA);
          add(ODIAG_JDOT, -(t[3] * sg) * dA);
          add(ODIAG_PHI, 0.5 * std::fabs(pc[BX1]) * sg * dA);
        }
    }
    if (iBX < 0) continue;
    const int act[3] = {B.n[0] > 1, B.n[1] > 1, B.n[2] > 1};
    const int ndim = act[0] + act[1] + act[2];
    const double *ub[3];
    for (int d = 0; d < 3; ++d) ub[d] = U[c].data() + B.off * nc + (int64_t)(iBX + d) * B.cells;
    for (int k = B.s[2]; k <= B.e[2]; ++k)
      for (int j = B.s[1]; j <= B.e[1]; ++j)
        for (int i = B.s[0]; i <= B.e[0]; ++i) {
          const int64_t id = B.idx(k, j, i);
          if (excised(B, k, j, i)) continue;
          for (int d = 0; d < ndim; ++d)
            out[ODIAG_UBMAX] = std::fmax(out[ODIAG_UBMAX], std::fabs(ub[d][id]));
        }
    if (ndim == 2) {
      const int k = B.s[2];
      const int64_t sj = B.n[0];
      for (int j = B.s[1] + 1; j <= B.e[1]; ++j)
        for (int i = B.s[0] + 1; i <= B.e[0]; ++i) {
          const int64_t id = B.idx(k, j, i);
          if (excised(B, k, j, i) || excised(B, k, j, i - 1) || excised(B, k, j - 1, i) ||
              excised(B, k, j - 1, i - 1))
            continue;
          const double *b1 = ub[0], *b2 = ub[1];
          const double d1 = (b1[id] + b1[id - sj] - b1[id - 1] - b1[id - sj - 1]) / (2 * B.dx[0]);
          const double d2 = (b2[id] + b2[id - 1] - b2[id - sj] - b2[id - sj - 1]) / (2 * B.dx[1]);
          out[ODIAG_DIVB] = std::fmax(out[ODIAG_DIVB], std::fabs(d1 + d2));
        }
    } else if (ndim == 3) {
      const int64_t st[3] = {1, B.n[0], (int64_t)B.n[0] * B.n[1]};
      for (int k = B.s[2] + 1; k <= B.e[2]; ++k)
        for (int j = B.s[1] + 1; j <= B.e[1]; ++j)
          for (int i = B.s[0] + 1; i <= B.e[0]; ++i) {
            const int64_t id = B.idx(k, j, i);
            bool touch = false;
            for (int c8 = 0; c8 < 8; ++c8)
              touch = touch || excised(B, k - (c8 >> 2), j - ((c8 >> 1) & 1), i - (c8 & 1));
            if (touch) continue;
            double tot = 0.0;
            for (int d = 0; d < 3; ++d) {
              double acc = 0.0;
              for (int dk = 0; dk < 2; ++dk)
                for (int dj = 0; dj < 2; ++dj)
                  for (int di = 0; di < 2; ++di) {
                    const int off = d == 0 ? di : (d == 1 ? dj : dk);
                    const double sign = off == 0 ? 1.0 : -1.0;
                    acc = acc + sign * ub[d][id - dk * st[2] - dj * st[1] - di * st[0]];
                  }
              tot = tot + acc / (4 * B.dx[d]);
            }
            out[ODIAG_DIVB] = std::fmax(out[ODIAG_DIVB], std::fabs(tot));
          }
    }
  }
}

}  // namespace oracle
