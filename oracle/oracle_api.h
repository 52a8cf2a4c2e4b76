/* C API of the CPU oracle (TEST INFRASTRUCTURE ONLY; loaded by tests/ and by
 * bench.py's cpu_baseline leg through ctypes). */
#ifndef PANGU_ORACLE_API_H
#define PANGU_ORACLE_API_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define ORACLE_MAX_REF 8

/* Whole-run configuration.  Field meanings follow PANGU's input file
 * (<core>, <metric>, <electron>, <parthenon/mesh>, <parthenon/time>). */
typedef struct oracle_cfg {
  /* mesh */
  int nx[3];          /* mesh cells per dim (1 = inactive) */
  int nb[3];          /* meshblocks per dim */
  int ng;             /* ghost layers on active dims */
  double xmin[3], xmax[3];
  int bc[3][2];       /* 0 periodic, 1 outflow, 2 reflecting */
  /* time */
  double dt_init, dt_factor;
  /* core */
  int metric;         /* 0 minkowski, 1 bl, 2 cks, 3 mks */
  double a, h;
  double gamma, cfl;
  int riemann;        /* 0 laxf, 1 hll */
  int limiter;        /* 0 mc, 1 ppm4 */
  int recovery;       /* 0 chainer, 1 kastaun (robuster), 2 1dvsq */
  int enable_B, enable_heating;
  double density_floor, energy_floor, density_floor_pow, energy_floor_pow;
  double sigma_max, lorentz_max;
  /* electron */
  int heat_model;
  double fel_constant, gamma_e, gamma_p, ratio_min, ratio_max, fel_0;
  int limit_kel, suppress_highb_heat, enforce_positive_dissipation;
  /* excision (CKS) */
  int excise;
  double r_excise, dexcise, pexcise;
  /* Parthenon semantics switches */
  int stage_init;      /* 0: new stage container zeroed, 1: copied from base */
  int exchange_after;  /* 0: after the update (reference order), 1: after recovery */
  int cache_geometry;  /* 1: memoise metric evaluations (identical values) */
  /* problem */
  int problem;         /* 1 Orszag-Tang (unit box), 2 Fishbone-Moncrief torus */
  double rin, rmax, kappa, perturbation, beta, aphi_cut;
  uint64_t seed;
  int nthreads;        /* OpenMP threads (0 = default) */
  /* static mesh refinement (<parthenon/static_refinementN>): region r asks
   * for blocks of level ref_level[r] wherever they overlap
   * [ref_min[r], ref_max[r]]; nref = 0: a uniform mesh */
  int nref;
  int ref_level[ORACLE_MAX_REF];
  double ref_min[ORACLE_MAX_REF][3], ref_max[ORACLE_MAX_REF][3];
  int magnetic_mad;    /* CKS torus field: 0 sane, 1 mad */
} oracle_cfg;

/* fields for oracle_sim_get / oracle_sim_set */
enum {
  ORF_PRIM = 0, ORF_CONS = 1, ORF_FLUX1 = 2, ORF_FLUX2 = 3, ORF_FLUX3 = 4,
  ORF_ALFVEN = 5, ORF_EMF = 6, ORF_FLAG = 7, ORF_MASK = 8, ORF_C2P_STATUS = 9,
  ORF_C2P_ITERS = 10, ORF_FLAG_RECOVERY = 11, ORF_DUDT = 12
};
/* tasks for oracle_sim_task (container indices: 0 base, 1 stage "1") */
enum {
  ORT_FIX_PRIMITIVE = 0, ORT_CONSERVATIVE = 1, ORT_FLUXES = 2, ORT_CT = 3,
  ORT_FLUX_DIV = 4, ORT_UPDATE = 5, ORT_SOURCE = 6, ORT_RECOVERY = 7,
  ORT_HEATING = 8, ORT_FIX_RECOVERY = 9, ORT_EXCHANGE = 10, ORT_BC = 11,
  ORT_TIMESTEP = 12, ORT_FLUX_CORRECTION = 13
};

void *oracle_sim_create(const oracle_cfg *cfg);
void oracle_sim_destroy(void *sim);
int oracle_sim_init(void *sim);
int oracle_sim_step(void *sim, int ncycles);
/* one task on container c0 (c1 = output container for update/source),
 * beta_dt for update/source; returns dt estimate for ORT_TIMESTEP in *out */
int oracle_sim_task(void *sim, int task, int c0, int c1, double beta_dt, double *out);
int64_t oracle_sim_field_size(void *sim, int field);
int oracle_sim_get(void *sim, int field, int container, void *out);
int oracle_sim_set(void *sim, int field, int container, const void *in);
double oracle_sim_time(void *sim);
double oracle_sim_dt(void *sim);
int oracle_sim_ncycle(void *sim);
int oracle_sim_block_geom(void *sim, int b, double *xg_dx_out /* 6 */, int *n_out /* 9 */);
int oracle_sim_nblocks(void *sim);
/* {level, lx1, lx2, lx3} of block b */
int oracle_sim_block_loc(void *sim, int b, int *out /* 4 */);

/* Diagnostics of container c (diagnostics.cpp): shell integrals at the
 * interior x1 cell index `shell` of the mesh, volume integrals, div B.
 * mag (optional) receives the sum of |contribution| of each summed entry. */
enum {
  ODIAG_MDOT = 0, ODIAG_EDOT = 1, ODIAG_JDOT = 2, ODIAG_PHI = 3, ODIAG_MASS = 4,
  ODIAG_ENERGY = 5, ODIAG_DIVB = 6, ODIAG_UBMAX = 7, ODIAG_N = 8
};
int oracle_sim_diagnostics(void *sim, int container, int shell, double *out /* ODIAG_N */,
                           double *mag /* ODIAG_N or NULL */);

/* Operation counts of the counting build (liboracle_count.so, counting.h):
 * out[task * ORACLE_NOPS + op] accumulates, per ORT_* task of run_stage,
 * the operations of kind op (add/sub, mul, div, sqrt, compare, fmin/fmax,
 * fabs, exp, log, pow, sin/cos) since the last reset.  Returns -1 in the
 * ordinary build. */
enum { ORACLE_NOPS = 11, ORACLE_NTASKS = 14 };
int oracle_op_counts(uint64_t *out /* ORACLE_NTASKS * ORACLE_NOPS */);
void oracle_op_reset(void);

/* planted-mistake switch (tests only): 0 = faithful */
void oracle_plant(int which);

/* pointwise functions for unit tests */
int oracle_metric_at(int kind, const double x[4], double a, double h, double gcov[16],
                     double gcon[16], double *gdet);
/* physical-coordinate metric CalculatePhysicalMetric (before the code-basis Jacobian) */
void oracle_physical_metric(int kind, const double x[4], double a, double h, double gcov[16]);
void oracle_christoffel(int kind, const double x[4], double a, double h, double conn[64]);
void oracle_state(const double prim[10], const double gcov[16], const double gcon[16],
                  double out[17]);
void oracle_flux(double gamma, const double prim[10], const double gcov[16],
                 const double gcon[16], double gdet, int dir, double flux[10]);
void oracle_fms(double gamma, const double prim[10], const double gcov[16],
                const double gcon[16], int dir, double out[2]);
void oracle_recon(int limiter, const double q[5], double out[2]);
int oracle_c2p(int scheme, const double U[9], double prim[9], double gamma,
               const double gcov[16], const double gcon[16], int *iters);
/* |f(mu)| of Kastaun's master function at mu for the container-form
 * conserved vector Ucont (rho, energy, momentum[3], field[3]); 0 on
 * success, 1 or 2 when the state has no invariants. */
int oracle_kastaun_residual(const double Ucont[8], double mu, double gamma, const double gcov[16],
                            const double gcon[16], double *res);
/* 1Dvsq residual (scheme_1d_vsq.h:91-117) at v^2 = vsq, entropy K, for a
 * container-form conserved vector; *scale = Q~^2 + vsq (B^2 + W)^2.
 * Returns 1 when D <= 0 or vsq is outside [0, 1). */
int oracle_vsq_residual(const double Ucont[8], double K, double vsq, double gamma, const double gcov[16],
                        const double gcon[16], double *res, double *scale);
/* cell = {rho, energy, b^2, advected entropy, recovered entropy, advected
 * electron entropy}: heating_model.h apply() and computeHeatingFraction() */
double oracle_heating(const oracle_cfg *cfg, const double cell[6]);
double oracle_heating_fraction(const oracle_cfg *cfg, const double cell[6]);

#ifdef __cplusplus
}
#endif
#endif
