// CPU oracle for the GRMHD per-RK-stage update path.
//
// TEST INFRASTRUCTURE ONLY. This is a plain CPU restatement of PANGU's
// per-stage algorithm (pangu/src/**) and of the Parthenon semantics it relies
// on.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
// may load it, and only as the checker (or the timed CPU baseline).  The
// product library (pangu_amd/) never links or calls it.
//
// Every arithmetic expression keeps the operand order of the reference so
// that results are reproducible bit for bit on any IEEE host built with
// -ffp-contract=off.  Each function cites the reference file:line it follows.
//
// Parity pins (what checks each part; see tests/test_oracle_*.py):
//   metric_*            analytic Kerr-Schild / BL / Minkowski closed forms,
//                       g*g^-1 = I, det vs analytic sqrt(-g), Minkowski exact
//   christoffel         analytic Schwarzschild/Kerr Gamma (FD floor ~1e-6),
//                       symmetry in the lower indices, Minkowski = 0
//   state / tmunu /     u.u = -1, b.u = 0, F^0(P) = U, T_{mu nu} symmetric
//   contravariant_flux
//   fast_ms_speed       cold unmagnetised limit (speeds = v), light-speed cap
//   recon_mc / ppm4     exact for constant and linear data, monotone at
//                       extrema, brute force on tiny cases
//   riemann (laxf/hll)  consistency F(P,P) = F(P); HLL = LAXF for c+ = c-
//   ct                  corner div B preserved to round-off
//   recovery schemes    P -> U -> P round trips (tolerance 1e-9),
//                       failure codes for D <= 0 / v >= 1
//   heating             closed forms of the heating fraction per model
//   fixers              floor values and masks on constructed states
//   driver              conservation of sum(U_rho) on periodic OT,
//                       FM torus near-equilibrium over one step
// Parity unpinned (no reference build or fixture exists; semantics inferred
// from Parthenon and restated): ghost exchange order, stage-container
// initialisation, dt control, flux divergence form, block coordinates.
// Also unpinned: the MAD factor ((r/rin) sin theta)^3 exp(-r/400) of the CKS
// torus potential (problems.cpp, init_torus_cks).  The divergence and beta
// checks hold for any potential, so they cannot see a wrong factor.
#pragma once

#include <cmath>
#include <cstdint>
#include <vector>

#ifdef ORACLE_COUNTING
#include "counting.h"
#endif

namespace oracle {

#ifdef ORACLE_COUNTING
typedef Real real;  // operation-counting build (counting.h)
#else
typedef double real;
#endif

// Slot order of a full primitive/conserved vector
// (pangu/src/initialization/variable_mnemonics.h:7-10).
enum Slot { RHO = 0, ENY, UX1, UX2, UX3, BX1, BX2, BX3, ENT, KEL, NPRIM };
constexpr int NPRIM_RECV = 9;  // recovery/constants.h:14

enum MetricKind { MET_MINKOWSKI = 0, MET_BL = 1, MET_CKS = 2, MET_MKS = 3 };
enum Limiter { LIM_MC = 0, LIM_PPM4 = 1 };
enum Riemann { RS_LAXF = 0, RS_HLL = 1 };
enum Recovery { REC_CHAINER = 0, REC_KASTAUN = 1, REC_1DVSQ = 2 };
enum HeatModel { HM_CONSTANT = 0, HM_HOWES, HM_KAWAZURA, HM_WERNER, HM_ROWAN, HM_SHARMA };
enum BoundaryKind { BC_PERIODIC = 0, BC_OUTFLOW = 1, BC_REFLECT = 2 };

// Recovery constants (recovery/constants.h:15-23).
constexpr int MAX_NEWT_ITER = 30;
constexpr double NEWT_TOL = 1e-10;
constexpr double MIN_NEWT_TOL = 1e-10;
constexpr int EXTRA_NEWT_ITER = 2;
constexpr double NEWT_TOL2 = 1.0e-15;
constexpr double W_TOO_BIG = 1.e20;
constexpr double UTSQ_TOO_BIG = 1.e20;
constexpr double FAIL_VAL = 1.e30;

// Floor-mask bits written by fix_primitive (diagnostic; the reference keeps none).
enum FloorBits { FL_RHO = 1, FL_ENY = 2, FL_GAMMA = 4, FL_SIGMA = 8, FL_EXCISE = 16 };

// Runtime parameters (package_registration.cc:21-57,158-167).
struct Params {
  int metric;
  double a, h;
  double gamma;
  int riemann, limiter, recovery;
  int enable_B, enable_heating;
  double cfl;
  double density_floor, energy_floor, density_floor_pow, energy_floor_pow;
  double sigma_max, lorentz_max;
  int heat_model;
  double fel_constant, gamma_e, gamma_p, ratio_min, ratio_max;
  int limit_kel, suppress_highb_heat, enforce_positive_dissipation;
  int excise;
  double r_excise, dexcise, pexcise;
};

// Planted-mistake switch for the tests that prove the pins catch plausible
// bugs (0 = faithful; see tests/test_oracle_pins.py).  Never set otherwise.
extern int g_plant;
enum Plant {
  PLANT_NONE = 0,
  PLANT_CONN_INDEX = 1,   // Christoffel: swap two derivative indices
  PLANT_SHIFT_SIGN = 2,   // state: wrong sign of the shift term in u^i
  PLANT_PPM_TERM = 3,     // PPM4: drop the outer stencil term
  PLANT_KASTAUN_QBAR = 4, // Kastaun: wrong sign of the b^2 term in qbar
  PLANT_CT_SIGN = 5,      // CT: wrong sign of the X1-face B3 flux
  PLANT_EMT_PRESSURE = 6, // T^mu_nu: drop the magnetic pressure
  PLANT_NO_FLUX_CORRECTION = 7,  // refined mesh: skip flux correction
  PLANT_PROLONG_SLOPE = 8,       // prolongation: +-1/2 instead of +-1/4 of the slope
  PLANT_CKS_JACOBIAN = 9,        // CKS torus: wrong sign of dx/dphi in the KS -> CKS Jacobian
  PLANT_CKS_CURL = 10,           // CKS torus field: wrong sign of B^1 = -dA2/dz
  PLANT_HEAT_HOWES = 11,         // Howes: electron/proton mass ratio inverted under the root
  PLANT_HEAT_KAWAZURA = 12,      // Kawazura: (beta/15)^(+1.4) instead of ^(-1.4)
  PLANT_HEAT_WERNER = 13,        // Werner: sigma/2 instead of sigma/5
  PLANT_HEAT_ROWAN = 14,         // Rowan: beta_max = 1/(2 sigma) instead of 1/(4 sigma)
  PLANT_HEAT_SHARMA = 15,        // Sharma: T_p/T_e under the root instead of T_e/T_p
  PLANT_BL_GTHTH = 16            // BL: g_thth = r^2 mu (Kerr's Sigma) instead of the reference's r^2/mu
};

// Kokkos::max / Kokkos::min semantics (ternary on operator<).
static inline real kmax(real a, real b) { return (a < b) ? b : a; }
static inline real kmin(real a, real b) { return (b < a) ? b : a; }

// ---- metric.cpp ----
void physical_coords(int kind, const real x[4], real y[4], real h);
void physical_metric(int kind, const real x[4], real g[4][4], real h, real a);
void code_metric(int kind, const real x[4], real g[4][4], real h, real a);
real determinant4(const real m[4][4]);
int invert4(const real m[4][4], real inv[4][4]);
// ComputeMetricAtLocation; returns invert4's status (0 = singular, gcon untouched).
int metric_at(int kind, const real x[4], real a, real h, real gcov[4][4],
              real gcon[4][4], real *gdet);
void christoffel(int kind, const real x[4], real a, real h, real conn[4][4][4]);

// ---- physics.cpp ----
struct State {
  real ucon[4], ucov[4], bcon[4], bcov[4], bsq;
};
void calc_state(const real prim[NPRIM], const real gcov[4][4], const real gcon[4][4], State &s);
void tmunu_row(real gamma, const real prim[NPRIM], const real gcov[4][4],
               const real gcon[4][4], int dir, real t[4]);
void contravariant_flux(real gamma, const real prim[NPRIM], const real gcov[4][4],
                        const real gcon[4][4], real gdet, int dir, real flux[NPRIM]);
void fast_ms_speed(real gamma, const real prim[NPRIM], const real gcov[4][4],
                   const real gcon[4][4], int dir, real *vmax, real *vmin);
void recon_mc(real qm2, real qm1, real q0, real qp1, real qp2, real *ql_ip1, real *qr_i);
void recon_ppm4(real qm2, real qm1, real q0, real qp1, real qp2, real *ql_ip1, real *qr_i);

struct HeatCell {
  real rho, energy, bsq, adv_total, rec_total, adv_e;
};
real heating_fraction(const Params &p, const HeatCell &c);
real heating_apply(const Params &p, const HeatCell &c);
real clamp_by_ratio(real rmin, real rmax, real total, real electron);

// ---- recovery.cpp ----
// Each returns the scheme's status code (0 = success) and the number of
// solver iterations in *iters.
int c2p_2d(real U[NPRIM_RECV], real prim[NPRIM_RECV], real gamma,
           const real gcov[4][4], const real gcon[4][4], int *iters);
int c2p_1dvsq(real U[NPRIM_RECV], real prim[NPRIM_RECV], real gamma,
              const real gcov[4][4], const real gcon[4][4], int *iters);
int c2p_1d(real U[NPRIM_RECV], real prim[NPRIM_RECV], real gamma,
           const real gcov[4][4], const real gcon[4][4], int *iters);
int c2p_kastaun(real U[NPRIM_RECV], real prim[NPRIM_RECV], real gamma,
                const real gcov[4][4], const real gcon[4][4], int *iters);
real kastaun_root(real mu, real D, real q, real bsq, real rsq, real rbsq, real gamma);
void vsq_residual(const real U[NPRIM_RECV], real K, real vsq, real gamma, const real gcov[4][4],
                  const real gcon[4][4], real *res, real *scale);

}  // namespace oracle
