// extern "C" surface of the CPU oracle (TEST INFRASTRUCTURE ONLY).
#include <cstring>

#include "oracle.h"
#include "oracle_api.h"
#include "oracle_sim.h"

using namespace oracle;

extern "C" {

void *oracle_sim_create(const oracle_cfg *cfg) {
  if (!cfg) return nullptr;
  for (int d = 0; d < 3; ++d)
    if (cfg->nb[d] <= 0 || cfg->nx[d] % cfg->nb[d] != 0) return nullptr;
  Sim *s = new Sim();
  s->setup(*cfg);
  return s;
}

void oracle_sim_destroy(void *sim) { delete (Sim *)sim; }

int oracle_sim_init(void *sim) { return ((Sim *)sim)->init_problem(); }

int oracle_sim_step(void *sim, int n) {
  Sim *s = (Sim *)sim;
  for (int i = 0; i < n; ++i) s->cycle();
  return 0;
}

int oracle_sim_task(void *sim, int task, int c0, int c1, double beta_dt, double *out) {
  Sim *s = (Sim *)sim;
  s->ensure_stage_container();
  switch (task) {
    case ORT_FIX_PRIMITIVE: s->fix_primitive(c0); break;
    case ORT_CONSERVATIVE: s->calc_conservative(c0); break;
    case ORT_FLUXES: s->fluxes(c0); break;
    case ORT_CT: s->ct(); break;
    case ORT_FLUX_DIV: s->flux_div(); break;
    case ORT_UPDATE: s->update(beta_dt, c1); break;
    case ORT_SOURCE: s->source(c1, beta_dt); break;
    case ORT_RECOVERY: s->recovery(c1); break;
    case ORT_HEATING: s->heating(c1); break;
    case ORT_FIX_RECOVERY: s->fix_recovery(c1); break;
    case ORT_EXCHANGE: s->exchange(c1); break;
    case ORT_BC: s->apply_bc(c1); break;
    case ORT_FLUX_CORRECTION: s->flux_correction(); break;
    case ORT_TIMESTEP:
      if (out) *out = s->estimate_dt();
      break;
    default: return -1;
  }
  return 0;
}

static void *field_ptr(Sim *s, int field, int c, int64_t *bytes) {
  const int64_t n = s->total_cells;
  switch (field) {
    case ORF_PRIM: *bytes = n * s->np * 8; return s->P[c].data();
    case ORF_CONS: *bytes = n * s->nc * 8; return s->U[c].data();
    case ORF_FLUX1: *bytes = n * s->nc * 8; return s->F[0].data();
    case ORF_FLUX2: *bytes = n * s->nc * 8; return s->F[1].data();
    case ORF_FLUX3: *bytes = n * s->nc * 8; return s->F[2].data();
    case ORF_ALFVEN: *bytes = n * 3 * 8; return s->alfven.data();
    case ORF_EMF: *bytes = n * 3 * 8; return s->emf.data();
    case ORF_FLAG: *bytes = n * 8; return s->flag.data();
    case ORF_MASK: *bytes = n; return s->mask.data();
    case ORF_C2P_STATUS: *bytes = n * 4; return s->c2p_status.data();
    case ORF_C2P_ITERS: *bytes = n * 4; return s->c2p_iters.data();
    case ORF_FLAG_RECOVERY: *bytes = n * 8; return s->flag_rec.data();
    case ORF_DUDT: *bytes = n * s->nc * 8; return s->dudt.data();
  }
  *bytes = 0;
  return nullptr;
}

int64_t oracle_sim_field_size(void *sim, int field) {
  int64_t b = 0;
  field_ptr((Sim *)sim, field, 0, &b);
  return b;
}

int oracle_sim_get(void *sim, int field, int c, void *out) {
  int64_t b = 0;
  void *p = field_ptr((Sim *)sim, field, c, &b);
  if (!p) return -1;
  std::memcpy(out, p, (size_t)b);
  return 0;
}

int oracle_sim_set(void *sim, int field, int c, const void *in) {
  int64_t b = 0;
  void *p = field_ptr((Sim *)sim, field, c, &b);
  if (!p) return -1;
  std::memcpy(p, in, (size_t)b);
  ((Sim *)sim)->stage1_ready = true;
  return 0;
}

double oracle_sim_time(void *sim) { return ((Sim *)sim)->time; }
double oracle_sim_dt(void *sim) { return ((Sim *)sim)->dt; }
int oracle_sim_ncycle(void *sim) { return ((Sim *)sim)->ncycle; }

int oracle_sim_diagnostics(void *sim, int c, int shell, double *out, double *mag) {
  if (c < 0 || c > 1) return 1;
  double m[ODIAG_N];
  ((Sim *)sim)->diagnostics(c, shell, out, m);
  if (mag)
    for (int q = 0; q < ODIAG_N; ++q) mag[q] = m[q];
  return 0;
}

int oracle_sim_nblocks(void *sim) { return ((Sim *)sim)->nblk; }

int oracle_sim_block_loc(void *sim, int b, int *out) {
  Sim *s = (Sim *)sim;
  if (b < 0 || b >= s->nblk) return -1;
  out[0] = s->blk[b].level;
  for (int d = 0; d < 3; ++d) out[1 + d] = s->blk[b].lx[d];
  return 0;
}

int oracle_sim_block_geom(void *sim, int b, double *xd, int *n) {
  Sim *s = (Sim *)sim;
  if (b < 0 || b >= s->nblk) return -1;
  const Block &B = s->blk[b];
  for (int d = 0; d < 3; ++d) {
    xd[d] = B.xg[d];
    xd[3 + d] = B.dx[d];
    n[d] = B.n[d];
    n[3 + d] = B.s[d];
    n[6 + d] = B.e[d];
  }
  return 0;
}

void oracle_plant(int which) { g_plant = which; }

#ifdef ORACLE_COUNTING
static_assert((int)OP_N == (int)ORACLE_NOPS && (int)ORACLE_NTASKS == 14, "operation kinds and tasks");
int oracle_op_counts(uint64_t *out) {
  for (int t = 0; t < ORACLE_NTASKS; ++t)
    for (int q = 0; q < ORACLE_NOPS; ++q) out[t * ORACLE_NOPS + q] = g_task_ops[t][q];
  return 0;
}
void oracle_op_reset(void) {
  for (int t = 0; t < ORACLE_NTASKS; ++t)
    for (int q = 0; q < ORACLE_NOPS; ++q) g_task_ops[t][q] = 0;
}
#else
int oracle_op_counts(uint64_t *) { return -1; }
void oracle_op_reset(void) {}
#endif

#ifndef ORACLE_COUNTING
// Per-cell helpers for the tests (double interface; not in the counting build).
int oracle_metric_at(int kind, const double x[4], double a, double h, double gcov[16],
                     double gcon[16], double *gdet) {
  real gc[4][4], gn[4][4];
  for (int i = 0; i < 16; ++i) gn[i / 4][i % 4] = 0.0;
  const int ok = metric_at(kind, x, a, h, gc, gn, gdet);
  std::memcpy(gcov, gc, sizeof(gc));
  std::memcpy(gcon, gn, sizeof(gn));
  return ok;
}

void oracle_physical_metric(int kind, const double x[4], double a, double h, double gcov[16]) {
  real g[4][4];
  physical_metric(kind, x, g, h, a);
  std::memcpy(gcov, g, sizeof(g));
}

void oracle_christoffel(int kind, const double x[4], double a, double h, double conn[64]) {
  real c[4][4][4];
  christoffel(kind, x, a, h, c);
  std::memcpy(conn, c, sizeof(c));
}

void oracle_state(const double prim[10], const double gcov[16], const double gcon[16],
                  double out[17]) {
  State s;
  calc_state(prim, (const real(*)[4])gcov, (const real(*)[4])gcon, s);
  for (int i = 0; i < 4; ++i) {
    out[i] = s.ucon[i];
    out[4 + i] = s.ucov[i];
    out[8 + i] = s.bcon[i];
    out[12 + i] = s.bcov[i];
  }
  out[16] = s.bsq;
}

void oracle_flux(double gamma, const double prim[10], const double gcov[16],
                 const double gcon[16], double gdet, int dir, double flux[10]) {
  contravariant_flux(gamma, prim, (const real(*)[4])gcov, (const real(*)[4])gcon, gdet, dir, flux);
}

void oracle_fms(double gamma, const double prim[10], const double gcov[16],
                const double gcon[16], int dir, double out[2]) {
  fast_ms_speed(gamma, prim, (const real(*)[4])gcov, (const real(*)[4])gcon, dir, &out[0],
                &out[1]);
}

void oracle_recon(int limiter, const double q[5], double out[2]) {
  if (limiter == LIM_MC)
    recon_mc(q[0], q[1], q[2], q[3], q[4], &out[0], &out[1]);
  else
    recon_ppm4(q[0], q[1], q[2], q[3], q[4], &out[0], &out[1]);
}

int oracle_c2p(int scheme, const double U[9], double prim[9], double gamma, const double gcov[16],
               const double gcon[16], int *iters) {
  real Uc[9];
  std::memcpy(Uc, U, sizeof(Uc));
  const real(*gc)[4] = (const real(*)[4])gcov;
  const real(*gn)[4] = (const real(*)[4])gcon;
  switch (scheme) {
    case 0: return c2p_2d(Uc, prim, gamma, gc, gn, iters);
    case 1: return c2p_1dvsq(Uc, prim, gamma, gc, gn, iters);
    case 2: return c2p_1d(Uc, prim, gamma, gc, gn, iters);
    case 3: return c2p_kastaun(Uc, prim, gamma, gc, gn, iters);
  }
  return -1;
}

// |f(mu)| of Kastaun's master function (scheme_kastaun.h:84-108) at a
// given mu, with the invariants computed from the conserved state as
// SchemeKastaun::invert does (scheme_kastaun.h:167-213).  Ucont is the
// conserved vector as a container holds it (sqrt(-g) times rho u^t,
// T^t_t + rho u^t, T^t_i, B^i), normalised here by alpha / sqrt(-g) with
// the energy slot holding U_E - U_rho (invertor.cc:149-174).  An
// independent check of a root a recovery accepted.
int oracle_kastaun_residual(const double Ucont[8], double mu, double gamma, const double gcov[16],
                            const double gcon[16], double *res) {
  const real(*g4)[4] = (const real(*)[4])gcov;
  const real(*gn)[4] = (const real(*)[4])gcon;
  const real sg = std::sqrt(std::fabs(determinant4(g4)));
  const real alpha = 1.0 / std::sqrt(-gn[0][0]);
  const real aos = alpha / sg;
  real U[9];
  U[RHO] = aos * Ucont[0];
  U[ENY] = aos * (Ucont[1] - Ucont[0]);
  for (int a = 0; a < 3; ++a) U[UX1 + a] = aos * Ucont[2 + a];
  for (int a = 0; a < 3; ++a) U[BX1 + a] = aos * Ucont[5 + a];
  const real g[6] = {g4[1][1], g4[1][2], g4[1][3], g4[2][2], g4[2][3], g4[3][3]};
  const real det = g[0] * g[3] * g[5] + 2.0 * g[1] * g[2] * g[4] - g[0] * g[4] * g[4] -
                   g[1] * g[1] * g[5] - g[2] * g[2] * g[3];
  if (!(det > 0.0)) return 1;
  const real gi[6] = {(g[3] * g[5] - g[4] * g[4]) / det, (g[2] * g[4] - g[1] * g[5]) / det,
                      (g[1] * g[4] - g[2] * g[3]) / det, (g[0] * g[5] - g[2] * g[2]) / det,
                      (g[1] * g[2] - g[0] * g[4]) / det, (g[0] * g[3] - g[1] * g[1]) / det};
  const real D = U[RHO];
  if (!(D > 0.0)) return 2;
  const real lapse = std::sqrt(-1.0 / gn[0][0]);
  const real Qcon0 = gn[0][0] * U[ENY] + gn[0][1] * U[UX1] + gn[0][2] * U[UX2] + gn[0][3] * U[UX3];
  const real tau = lapse * Qcon0 - D;
  const real S[3] = {U[UX1], U[UX2], U[UX3]};
  const real Su[3] = {gi[0] * S[0] + gi[1] * S[1] + gi[2] * S[2], gi[1] * S[0] + gi[3] * S[1] + gi[4] * S[2],
                      gi[2] * S[0] + gi[4] * S[1] + gi[5] * S[2]};
  const real B[3] = {U[BX1], U[BX2], U[BX3]};
  const real Bd[3] = {g[0] * B[0] + g[1] * B[1] + g[2] * B[2], g[1] * B[0] + g[3] * B[1] + g[4] * B[2],
                      g[2] * B[0] + g[4] * B[1] + g[5] * B[2]};
  const real Bsq = B[0] * Bd[0] + B[1] * Bd[1] + B[2] * Bd[2];
  const real Ssq = Su[0] * S[0] + Su[1] * S[1] + Su[2] * S[2];
  const real SdotB = B[0] * S[0] + B[1] * S[1] + B[2] * S[2];
  const real bsq = Bsq / D, rsq = Ssq / (D * D), rbsq = SdotB * SdotB / (D * D * D), q = tau / D;
  *res = std::fabs(kastaun_root(mu, D, q, bsq, rsq, rbsq, gamma));
  return 0;
}

// The 1Dvsq residual (func_1d_vsq, scheme_1d_vsq.h:91-117) at v^2 = vsq
// for a container-form conserved vector, normalised as above, with the
// entropy K the scheme holds fixed; res and scale as vsq_residual.
int oracle_vsq_residual(const double Ucont[8], double K, double vsq, double gamma, const double gcov[16],
                        const double gcon[16], double *res, double *scale) {
  const real(*g4)[4] = (const real(*)[4])gcov;
  const real(*gn)[4] = (const real(*)[4])gcon;
  const real sg = std::sqrt(std::fabs(determinant4(g4)));
  const real alpha = 1.0 / std::sqrt(-gn[0][0]);
  const real aos = alpha / sg;
  real U[NPRIM_RECV] = {0};
  U[RHO] = aos * Ucont[0];
  U[ENY] = aos * (Ucont[1] - Ucont[0]);
  for (int a = 0; a < 3; ++a) U[UX1 + a] = aos * Ucont[2 + a];
  for (int a = 0; a < 3; ++a) U[BX1 + a] = aos * Ucont[5 + a];
  if (!(U[RHO] > 0.0) || !(vsq >= 0.0 && vsq < 1.0)) return 1;
  vsq_residual(U, K, vsq, gamma, g4, gn, res, scale);
  return 0;
}

static Params heat_params(const oracle_cfg *c) {
  Params p{};
  p.gamma = c->gamma;
  p.heat_model = c->heat_model;
  p.fel_constant = c->fel_constant;
  p.gamma_e = c->gamma_e;
  p.gamma_p = c->gamma_p;
  p.ratio_min = c->ratio_min;
  p.ratio_max = c->ratio_max;
  p.limit_kel = c->limit_kel;
  p.suppress_highb_heat = c->suppress_highb_heat;
  p.enforce_positive_dissipation = c->enforce_positive_dissipation;
  return p;
}

double oracle_heating(const oracle_cfg *c, const double cell[6]) {
  HeatCell hc{cell[0], cell[1], cell[2], cell[3], cell[4], cell[5]};
  return heating_apply(heat_params(c), hc);
}

double oracle_heating_fraction(const oracle_cfg *c, const double cell[6]) {
  HeatCell hc{cell[0], cell[1], cell[2], cell[3], cell[4], cell[5]};
  return heating_fraction(heat_params(c), hc);
}

#endif  // ORACLE_COUNTING

}  // extern "C"
