/*
 * pangu_sim: host driver above the pangu_hip C ABI.
 *
 * It restates what Parthenon does around PANGU's task list on one GPU: the
 * uniform block mesh, the "base"/"1"/"dUdt" stage containers, the RK2 stage
 * sequence of Simulator::MakeTaskCollection (pangu/src/task_list/
 * ideal_grmhd.cc:23-120), the initial data of the FM torus / Orszag-Tang
 * problems and dt control.  Field arrays live in device memory, one arena
 * per field laid out [block][component][k][j][i] over ghosted blocks.
 * Entry points with no cited counterpart are the driver's own (field
 * access, timing, tuning); a Parthenon build needs none of them.
 */
#ifndef PANGU_SIM_H
#define PANGU_SIM_H
#include <stdint.h>

#include "pangu_hip.h"
#include "pangu_mlp.h"

#ifdef __cplusplus
extern "C" {
#endif

#define PANGU_SIM_MAX_REF 8

/* Run configuration: mesh, time, physics, problem.  Field meanings follow
 * PANGU's input file. */
typedef struct pangu_sim_cfg {
  int nx[3];
  int nb[3];
  int ng;
  double xmin[3], xmax[3];
  int bc[3][2];
  double dt_init, dt_factor;
  int metric;
  double a, h;
  double gamma, cfl;
  int riemann, limiter, recovery;
  int enable_B, enable_heating;
  double density_floor, energy_floor, density_floor_pow, energy_floor_pow;
  double sigma_max, lorentz_max;
  int heat_model;
  double fel_constant, gamma_e, gamma_p, ratio_min, ratio_max, fel_0;
  int limit_kel, suppress_highb_heat, enforce_positive_dissipation;
  int excise;
  double r_excise, dexcise, pexcise;
  int stage_init;      /* 0: stage container starts zeroed, 1: copied from base */
  int exchange_after;  /* 0: exchange after the update (reference order), 1: after recovery */
  int cache_geometry;  /* unused by the device path (tables are always used) */
  int problem;         /* 1 Orszag-Tang (unit box), 2 FM torus */
  double rin, rmax, kappa, perturbation, beta, aphi_cut;
  uint64_t seed;
  int nthreads;
  /* static mesh refinement (<parthenon/static_refinementN>): region r asks
   * for blocks of level ref_level[r] (1 .. 20) wherever they overlap
   * [ref_min[r], ref_max[r]]; neighbours are refined so that adjacent blocks
   * differ by at most one level.  nref = 0: a uniform mesh. */
  int nref;
  int ref_level[PANGU_SIM_MAX_REF];
  double ref_min[PANGU_SIM_MAX_REF][3], ref_max[PANGU_SIM_MAX_REF][3];
  int magnetic_mad;    /* CKS torus field: 0 sane, 1 mad (<fm_torus>/magnetic_field) */
} pangu_sim_cfg;

enum { PANGU_SIM_TASKS = 0, PANGU_SIM_FUSED = 1, PANGU_SIM_GRAPH = 2 };

typedef struct pangu_sim pangu_sim;

int pangu_sim_create(const pangu_sim_cfg *cfg, int device, pangu_sim **out);
/* An ensemble of independent meshes in one context (pangu_create_batch):
 * every kernel launch covers all of them, each mesh keeps its own dt.  The
 * meshes must share the physics options, dt_factor, stage_init and
 * exchange_after; mesh m's blocks follow mesh m-1's in the arenas.  Task
 * mode (host dt) needs a single mesh. */
int pangu_sim_create_batch(const pangu_sim_cfg *cfgs, int nmesh, int device, pangu_sim **out);
/* dt, time, first block and block count of mesh m. */
int pangu_sim_mesh_state(pangu_sim *s, int m, double *dt, double *time, int *first_block,
                         int *nblocks);
int pangu_sim_nblocks(pangu_sim *s);
/* pangu_diagnostics on container 0 (base) or 1 (stage "1"), one shell index
 * per mesh; out (host) receives nmesh * PANGU_NDIAG doubles. */
int pangu_sim_diagnostics(pangu_sim *s, int container, const int *shell, double *out);
/* The simulation's C-ABI context, the handles of its containers 0 (base)
 * and 1 (stage "1") and its stream, for callers that drive pangu_hip
 * entry points on the simulation's fields directly. */
int pangu_sim_context(pangu_sim *s, pangu_ctx **ctx, int *handles, void **stream);
int pangu_sim_destroy(pangu_sim *s);
/* Writes the initial data of cfg.problem into the base container:
 * ProblemGenerator of pangu/problem/orszag_tang_vortex/problem_generator.cpp:16-69
 * (problem 1), fm_torus_mks/problem_generator.cpp:103-289 (2) and
 * fm_torus_cks/problem_generator.cpp:143-334 (3), then fills the base
 * container's ghost cells (exchange, then physical boundaries). */
int pangu_sim_init(pangu_sim *s);
/* Runs ncycles RK2 cycles in the given mode (tasks: one C-ABI call per
 * reference task with a host-side dt; fused: pangu_stage with device dt;
 * graph: one captured cycle replayed).  Returns when the work is enqueued;
 * pangu_sim_sync waits. */
int pangu_sim_step(pangu_sim *s, int ncycles, int mode);
int pangu_sim_sync(pangu_sim *s);
/* Runs ncycles and returns the elapsed milliseconds between HIP events
 * recorded on the simulation stream around them. */
int pangu_sim_timed_cycles(pangu_sim *s, int ncycles, int mode, double *ms);
/* Times each kernel phase of one fused stage on the simulation stream:
 * ms_out[n] = mean milliseconds over reps of phase n (see pangu_sim_phase_name). */
int pangu_sim_time_phases(pangu_sim *s, int reps, double *ms_out, int max_phases);
const char *pangu_sim_phase_name(int n);
/* Runs ncycles fused cycles with a HIP event recorded on the simulation
 * stream after every kernel launch.  Returns the number of distinct kernels
 * n; ms_out[i] is kernel i's total milliseconds per cycle and calls_out[i]
 * its launches per cycle (names from pangu_sim_kernel_name). */
int pangu_sim_time_kernels(pangu_sim *s, int ncycles, double *ms_out, int *calls_out,
                           int max_kernels);
const char *pangu_sim_kernel_name(pangu_sim *s, int n);
/* The loop a caller would write around the one-mesh API: ncycles cycles,
 * each running every simulation's cycle in turn (mode fused or graph), all
 * enqueued on the first simulation's stream.  ms = elapsed time between
 * HIP events around the loop. */
int pangu_sim_plain_loop(pangu_sim **sims, int n, int ncycles, int mode, double *ms);
/* Autotuner: whether the X1 sweep shares reconstructions, then for each
 * kernel family the faster tile order, then for each flux sweep the
 * workgroup shape (one lane per face, or two: pangu_set_pair_flux), then
 * the faces per thread of the X2 and X3 sweeps (0 or 2, 3, 4, 6, 8, 12:
 * pangu_set_flux_march), then the cells per tile of each family (256, 128
 * or 64: pangu_set_tile_size), then whether recovery runs as a persistent
 * work-queue kernel (pangu_set_persistent_recovery), then whether to split
 * the source kernel,
 * each candidate timed with the kernel clock over `cycles` fused cycles.
 * Every field, dt and time is restored afterwards, so tuning does not
 * advance the run.  Results are cached per block-shape class: with
 * use_cache != 0 a result stored for the same block shapes and options in
 * this process is reused without measuring.  out_orders (PANGU_KF_COUNT
 * ints), out_split, out_share, out_pair (direction mask), out_march (2
 * ints: X2, X3), out_tiles (PANGU_KF_COUNT ints) and out_persist may be
 * NULL. */
int pangu_sim_autotune(pangu_sim *s, int cycles, int use_cache, int *out_orders, int *out_split,
                       int *out_share, int *out_pair, int *out_march, int *out_tiles,
                       int *out_persist);
/* pangu_set_persistent_recovery on the simulation's context (drops the graph). */
int pangu_sim_set_persistent_recovery(pangu_sim *s, int on);
/* pangu_set_tile_size on the simulation's context (drops the graph). */
int pangu_sim_set_tile_size(pangu_sim *s, int kernel_family, int cells);
/* pangu_set_flux_march on the simulation's context (drops the graph). */
int pangu_sim_set_flux_march(pangu_sim *s, int dir, int depth);
/* pangu_set_c2p_guess on the simulation's context (drops the graph). */
int pangu_sim_set_c2p_guess(pangu_sim *s, pangu_mlp *m, double delta);
/* pangu_set_c2p_gate on the simulation's context (re-captures the cycle). */
int pangu_sim_set_c2p_gate(pangu_sim *s, int min_prev_iters);
/* on != 0: the Kastaun recovery records each cell's root (field 13,
 * c2p_root: mu where it succeeded, NaN where it failed); 0 stops it. */
int pangu_sim_record_roots(pangu_sim *s, int on);
/* Sets a kernel family's tile order (pangu_set_tile_order) and drops the
 * captured graph so the next graph-mode step captures the new launches. */
int pangu_sim_set_tile_order(pangu_sim *s, int kernel_family, int order);
/* pangu_set_split_source on the simulation's context (drops the graph). */
int pangu_sim_set_split_source(pangu_sim *s, int split);
/* pangu_set_fuse_ct on the simulation's context (drops the graph). */
int pangu_sim_set_fuse_ct(pangu_sim *s, int fuse);
/* pangu_set_overlap_exchange on the simulation's context (drops the graph). */
int pangu_sim_set_overlap_exchange(pangu_sim *s, int on);
/* pangu_set_fuse_update_recovery on the simulation's context (drops the graph). */
int pangu_sim_set_fuse_update_recovery(pangu_sim *s, int on);
/* pangu_set_fuse_fix_x1 on the simulation's context (drops the graph). */
int pangu_sim_set_fuse_fix_x1(pangu_sim *s, int on);
/* pangu_set_share_reconstruction on the simulation's context (drops the graph). */
int pangu_sim_set_share_reconstruction(pangu_sim *s, int share);
int pangu_sim_set_pair_flux(pangu_sim *s, int dir_mask);
/* One task on container c0 (c1 output of update/source), host beta*dt. */
int pangu_sim_task(pangu_sim *s, int task, int c0, int c1, double beta_dt, double *out);
int64_t pangu_sim_field_size(pangu_sim *s, int field);
int pangu_sim_get(pangu_sim *s, int field, int container, void *out);
int pangu_sim_set(pangu_sim *s, int field, int container, const void *in);
double pangu_sim_time(pangu_sim *s);
double pangu_sim_dt(pangu_sim *s);
int pangu_sim_block_geom(pangu_sim *s, int b, double *xg_dx_out, int *n_out);
/* Refinement level and logical location of block b: out = {level, lx1, lx2, lx3}. */
int pangu_sim_block_loc(pangu_sim *s, int b, int *out);
const char *pangu_sim_last_error(pangu_sim *s);
/* Times a plain copy kernel of n doubles (mean ms over reps); used to
 * calibrate hardware byte counters against a kernel of known traffic. */
int pangu_sim_copy_calibration(pangu_sim *s, int64_t n, int reps, double *ms);

#ifdef __cplusplus
}
#endif
#endif
