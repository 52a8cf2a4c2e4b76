/*
 * pangu_hip: MI355X (gfx950) FP64 GRMHD per-RK-stage update path.
 *
 * C ABI replacing the per-stage task functions that PANGU's task list
 * schedules (pangu/src/task_list/ideal_grmhd.cc:62-94,104-110).  A caller
 * (the Parthenon task list, or this package's own host driver) owns every
 * field array; the library owns geometry tables, scratch and launch state.
 *
 * Conventions
 *  - All field pointers are device pointers to FP64 arrays of one meshblock,
 *    laid out (k, j, i) with i fastest over the ghosted extent, exactly what
 *    Parthenon's VariablePack(b, l).data() exposes.
 *  - Components are packed in PANGU's registration order
 *    (package_registration.cc:84-96): density, energy, weighted_velocity[3],
 *    [magnetic_field[3]], entropy, [electron_entropy].
 *  - Every function returns 0 on success or a negative PANGU_E* code;
 *    pangu_last_error() gives the text.  Nothing throws across the ABI.
 *  - `stream` is a hipStream_t (NULL = default stream).  Work is
 *    stream-ordered and asynchronous unless stated otherwise.
 */
#ifndef PANGU_HIP_H
#define PANGU_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  PANGU_OK = 0,
  PANGU_EINVAL = -1,   /* invalid argument or configuration */
  PANGU_EDEVICE = -2,  /* HIP runtime error */
  PANGU_ENOMEM = -3,   /* device allocation failed */
  PANGU_ESTATE = -4    /* call out of order (e.g. no container registered) */
};

enum { PANGU_METRIC_MINKOWSKI = 0, PANGU_METRIC_BL = 1, PANGU_METRIC_CKS = 2, PANGU_METRIC_MKS = 3 };
enum { PANGU_RIEMANN_LAXF = 0, PANGU_RIEMANN_HLL = 1 };
enum { PANGU_LIMITER_MC = 0, PANGU_LIMITER_PPM4 = 1 };
/* chainer = 2D -> 1Dvsq -> 1D Newton with flag + FixRecovery; robuster =
 * Kastaun false position; onedvsq = the 1Dvsq solver alone, with flag. */
enum { PANGU_RECOVERY_CHAINER = 0, PANGU_RECOVERY_ROBUSTER = 1, PANGU_RECOVERY_ONEDVSQ = 2 };
enum { PANGU_HEAT_CONSTANT = 0, PANGU_HEAT_HOWES, PANGU_HEAT_KAWAZURA, PANGU_HEAT_WERNER,
       PANGU_HEAT_ROWAN, PANGU_HEAT_SHARMA };
enum { PANGU_BC_PERIODIC = 0, PANGU_BC_OUTFLOW = 1, PANGU_BC_REFLECT = 2 };

/* Physics parameters: the <core>, <metric> and <electron> blocks of PANGU's
 * input file (package_registration.cc:21-57,158-167). */
typedef struct pangu_config {
  int metric;
  double a, h;
  double gamma;  /* adiabatic_index */
  double cfl;
  int riemann, limiter, recovery;
  int enable_B, enable_heating;
  double density_floor, energy_floor, density_floor_pow, energy_floor_pow;
  double sigma_max, lorentz_max;
  int heat_model;
  double fel_constant, gamma_e, gamma_p, ratio_min, ratio_max;
  int limit_kel, suppress_highb_heat, enforce_positive_dissipation;
  int excise;
  double r_excise, dexcise, pexcise;
  /* mesh topology for ghost exchange and physical boundaries */
  int nb[3];     /* meshblocks per dim of the root grid (refinement level 0) */
  int bc[3][2];  /* PANGU_BC_* for inner/outer faces of each dim */
} pangu_config;

/* One meshblock: Parthenon UniformCartesian coordinates,
 * Xc(idx) = xmin - ng*dx + (idx + 0.5)*dx, Xf(idx) = xmin - ng*dx + idx*dx.
 * With static mesh refinement (Parthenon LogicalLocation), `level` counts
 * refinements of the root grid and lx is the location among the
 * nb[d] << level blocks of that level on active dims.  The blocks of a mesh
 * must tile its domain, share one shape, and neighbour (faces, edges and
 * corners) only blocks at most one level apart. */
typedef struct pangu_block {
  int nx[3];      /* interior cells (1 on inactive dims) */
  int ng[3];      /* ghost layers (0 on inactive dims) */
  int lx[3];      /* logical location in the block grid */
  double xmin[3]; /* interior lower corner */
  double dx[3];   /* cell widths */
  int level;      /* refinement level (0 = root grid) */
} pangu_block;

/* A stage container (Parthenon MeshData): per block, per component device
 * pointers.  prim and cons are [nblocks * ncomp] tables; flux[d] is
 * [nblocks * ncons] or NULL on inactive directions. */
typedef struct pangu_container {
  double *const *prim;
  double *const *cons;
  double *const *flux[3];
} pangu_container;

/* Fields shared by all containers (Metadata::OneCopy) and optional
 * per-cell diagnostics.  Any table may be NULL when the configuration does
 * not use it (emf without B, flag for robuster). */
typedef struct pangu_shared {
  double *const *alfven;   /* [nblocks * 3] face wave speeds */
  double *const *emf;      /* [nblocks * 3] corner EMFs */
  double *const *flag;     /* [nblocks] 1.0 ok / 0.0 recovery failed */
  uint8_t *const *mask;    /* [nblocks] floor bits of the last fix_primitive */
  int32_t *const *c2p_status; /* [nblocks] recovery status code (interior) */
  int32_t *const *c2p_iters;  /* [nblocks] recovery iterations (interior) */
  double *const *flag_recovery; /* [nblocks] flag as left by recovery, before fix_recovery */
  double *const *c2p_root;    /* [nblocks] Kastaun root mu of the last recovery where it
                                 succeeded, NaN where it failed (interior; robuster only) */
} pangu_shared;

typedef struct pangu_ctx pangu_ctx;

/* Creates a context for one block partition: validates the configuration
 * (replaces the PARTHENON_REQUIRE_THROWS / std::invalid_argument checks of
 * package_registration.cc:38-40, riemann_solver.cc:25, heating_model.h:118),
 * builds the FP64 geometry tables on the host and uploads them. */
int pangu_create(const pangu_config *cfg, const pangu_block *blocks, int nblocks, int device,
                 pangu_ctx **out);
/* Batched form (no reference counterpart): one context over several
 * independent meshes, e.g. an ensemble of tori plus 3-D boxes, whose blocks
 * may differ in shape between meshes.  cfgs[m] configures mesh m; they must
 * agree on every physics parameter except the metric's a and h, the block
 * topology nb and the boundaries bc.  Blocks are grouped by mesh:
 * block_mesh[b] is non-decreasing from 0 to nmesh - 1, and the blocks of one
 * mesh share a shape.  Each kernel launch then covers every block of every
 * mesh; ghost exchange stays inside a mesh and each mesh keeps its own dt.
 * pangu_create(cfg, blocks, n, dev, out) is the one-mesh case. */
int pangu_create_batch(const pangu_config *cfgs, int nmesh, const pangu_block *blocks,
                       const int *block_mesh, int nblocks, int device, pangu_ctx **out);
/* Frees what pangu_create allocated (geometry tables, pointer tables,
 * refinement tables); never the caller's arrays.  No reference counterpart:
 * Parthenon owns the device fields there. */
int pangu_destroy(pangu_ctx *ctx);
/* The message of the last failed call on ctx (no reference counterpart). */
const char *pangu_last_error(const pangu_ctx *ctx);

/* Host-only: one entry of the geometry tables at code coordinates x
 * (loc 0 = cell centre, 1..3 = lower face normal to x1..x3).  out receives
 * 102 (centre) or 24 (face) doubles in the table component order: gcov[16],
 * gcon[16], sqrt|g|, 1/sqrt(-g^00), sqrt(-1/g^00), the two floors, r and
 * Gamma[64] for centres; gcov[16], g^0mu[4], g^dd, sqrt|g|, 1/sqrt(-g^00),
 * valid for faces.  Replaces ComputeMetricAtLocation/ComputeChristoffelConnection
 * (metric/christoffel.h:53-148) as evaluated at table build time. */
int pangu_geometry_entry(const pangu_config *cfg, int loc, const double x[4], double *out);

/* Registers a container's pointer tables (copied to the device once);
 * returns a handle used by the task functions.  Pointers must stay valid.
 * Stands in for the variable packs every reference task builds on entry
 * (md->PackVariables / PackVariablesAndFluxes, riemann_solver/hll.cc:37,49):
 * the packs are resolved once here instead of per call. */
int pangu_register_container(pangu_ctx *ctx, const pangu_container *c, int *handle);
/* Registers the tables shared by all containers (wave speeds, EMFs, flags,
 * masks, recovery codes): the "alfven" pack of riemann_solver/hll.cc:51-53
 * and the other package-level fields the reference tasks pack by name. */
int pangu_register_shared(pangu_ctx *ctx, const pangu_shared *s);  /* again: replaces the tables */

/* ---- task functions (one per reference task) ---- */
/* FixPrimitive (fixer/primitive_fixer.cc:15) */
int pangu_fix_primitive(pangu_ctx *ctx, int md, void *stream);
/* CalculateConservative (riemann_solver/conservative.cc:15) */
int pangu_calculate_conservative(pangu_ctx *ctx, int md, void *stream);
/* CalculateFluxes -> CalculateLAXF / CalculateHLL (riemann_solver/riemann_solver.cc:13) */
int pangu_calculate_fluxes(pangu_ctx *ctx, int md, void *stream);
/* ConstraintedTransport (constrained_transport/constrained_transport.cc:13) */
int pangu_constrained_transport(pangu_ctx *ctx, int md, void *stream);
/* Parthenon FluxDivergence(mc0, mdudt) (task_list/ideal_grmhd.cc:73) */
int pangu_flux_divergence(pangu_ctx *ctx, int md_in, int md_dudt, void *stream);
/* Parthenon UpdateIndependentData(mc_init, mdudt, beta*dt, mc1) (ideal_grmhd.cc:75) */
int pangu_update_independent(pangu_ctx *ctx, int md_init, int md_dudt, double beta_dt, int md_out,
                             void *stream);
/* AddGeometricSource(md, dt) (riemann_solver/source_term.cc:15) */
int pangu_add_geometric_source(pangu_ctx *ctx, int md, double beta_dt, void *stream);
/* Recovery (recovery/invertor.cc:18) */
int pangu_recovery(pangu_ctx *ctx, int md, void *stream);
/* ApplyElectronHeating (riemann_solver/electron_heating.cc:14) */
int pangu_apply_electron_heating(pangu_ctx *ctx, int md, void *stream);
/* FixRecovery (fixer/recovery_fixer.cc:12), serial lexicographic semantics */
int pangu_fix_recovery(pangu_ctx *ctx, int md, void *stream);
/* EstimateTimestepMesh (initialization/timestep_estimation.cc:13); synchronous
 * like the reference's par_reduce: writes the estimate of mesh m to
 * dt_out[m] (one value for a one-mesh context). */
int pangu_estimate_timestep(pangu_ctx *ctx, int md, double *dt_out, void *stream);

/* ---- Parthenon-side pieces the path needs on one GPU ---- */
/* Ghost exchange (face/edge/corner neighbours, periodic wrap) of the fields
 * read in ghost zones: primitives and flag (AddBoundaryExchangeTasks,
 * ideal_grmhd.cc:78).  Across refinement levels a ghost cell covered by a
 * finer block takes the mean of the 2^dim cells it covers (the flag: their
 * minimum); one covered by a coarser block takes the coarse value plus
 * minmod-limited slopes times +-1/4 of a coarse cell per active dim (the
 * flag: the coarse value). */
int pangu_exchange_ghosts(pangu_ctx *ctx, int md, void *stream);
/* Flux correction at refinement boundaries (parthenon::AddFluxCorrectionTasks,
 * ideal_grmhd.cc:71-72): every face flux of a coarse block on a face shared
 * with finer blocks becomes the mean of the 2^(dim-1) fine face fluxes that
 * cover it, for all conserved components.  Runs after CalculateFluxes and
 * ConstraintedTransport, before FluxDivergence; a no-op on a single-level
 * mesh. */
int pangu_flux_correction(pangu_ctx *ctx, int md, void *stream);
/* Physical outflow / reflecting boundaries (ApplyBoundaryConditionsMD, ideal_grmhd.cc:104). */
int pangu_apply_boundary_conditions(pangu_ctx *ctx, int md, void *stream);

/* ---- fused stage (performance path; no reference counterpart) ----
 * Runs region 1 + region 2 of one RK stage (ideal_grmhd.cc:52-116) with
 * fused kernels: fix+cons, three flux sweeps, CT, flux correction (on a
 * refined mesh), flux-div+update+source, exchange, recovery+heating,
 * fix_recovery and boundaries.  dt_state is a
 * device array {dt, time} per mesh: the stage reads each mesh's beta*dt
 * from it, and when last_stage != 0 it also estimates the timestep and
 * applies the dt control time += dt, dt = min(dt_factor*dt, estimate) in
 * place.  Nothing is read
 * back to the host, so a whole cycle can be captured in a hipGraph and
 * replayed.  exchange_after: 0 = reference order (after the update),
 * 1 = after recovery. */
int pangu_stage(pangu_ctx *ctx, int md_init, int md0, int md1, double beta, double *dt_state,
                double dt_factor, int last_stage, int exchange_after, void *stream);

/* ---- launch tuning (no reference counterpart) ----
 * These options change how kernels are launched, never what they compute.
 * Tile order of each kernel family: plane-major (consecutive workgroups
 * sweep one k plane) or column-major (consecutive workgroups walk k over one
 * 256-cell run of the (j, i) plane, reusing x3-independent geometry from L2).
 * Results do not depend on it.  pangu_create sets the order measured best
 * on gfx950 for each family.  Takes effect on the next launch (re-capture
 * graphs). */
enum {
  PANGU_KF_FIX_CONS = 0, PANGU_KF_FLUX_X1, PANGU_KF_FLUX_X2, PANGU_KF_FLUX_X3, PANGU_KF_CT,
  PANGU_KF_DIV_UPDATE, PANGU_KF_RECOVERY, PANGU_KF_COUNT
};
enum { PANGU_ORDER_PLANE = 0, PANGU_ORDER_COLUMN = 1 };
int pangu_set_tile_order(pangu_ctx *ctx, int kernel_family, int order);
int pangu_get_tile_order(const pangu_ctx *ctx, int kernel_family);
/* Cells per tile (threads per workgroup) of a kernel family's launches:
 * 256 (default), 128 or 64.  Smaller tiles waste fewer lanes on planes of
 * small blocks (a 10x10 face plane fills 100 of 256 lanes).  The flux
 * sweeps use it in their one-lane-per-face form only.  Results are
 * identical. */
int pangu_set_tile_size(pangu_ctx *ctx, int kernel_family, int cells);
int pangu_get_tile_size(const pangu_ctx *ctx, int kernel_family);
/* Fused stage only: split = 1 runs the geometric source in a kernel of its
 * own after the flux divergence + update kernel (fewer registers per
 * kernel, one more pass over U); 0 fuses all three.  Results are identical. */
int pangu_set_split_source(pangu_ctx *ctx, int split);
/* share = 1: the X1 sweep reconstructs each cell once and hands the right
 * interface value to the next face through a wavefront shuffle, instead of
 * reconstructing both neighbours per face.  Results are identical. */
int pangu_set_share_reconstruction(pangu_ctx *ctx, int share);
/* Bit d-1 of dir_mask set: the flux sweep along d runs two lanes per face,
 * one per side (half the registers per lane, so more wavefronts in flight;
 * the two lanes share the stencil loads and exchange their wave speeds and
 * flux components).  Results are identical.  Ignored by an X1 sweep that
 * shares reconstructions (pangu_set_share_reconstruction). */
int pangu_set_pair_flux(pangu_ctx *ctx, int dir_mask);
/* depth > 0: the flux sweep along dir (2 or 3) runs `depth` consecutive
 * faces per thread along dir, reconstructing each cell of the column once
 * (depth + 1 reconstructions instead of 2 depth).  0 (default): one face
 * per thread.  Takes precedence over pangu_set_pair_flux for that
 * direction.  Results are identical.  Tile order (pangu_set_tile_order of
 * the sweep's family) applies to its chunks of faces. */
int pangu_set_flux_march(pangu_ctx *ctx, int dir, int depth);
/* on = 1: recovery (and heating) runs as a persistent kernel whose
 * workgroups, no more than the GPU holds at once, pull tiles from a queue
 * across every block of every mesh until none is left (the learned bracket
 * keeps its own kernel).  0 (default): one workgroup per tile.  Results
 * are identical. */
int pangu_set_persistent_recovery(pangu_ctx *ctx, int on);
/* fuse = 1 (default): the fused stage runs no constrained-transport kernel;
 * the flux-divergence + update kernel forms the edge EMFs from the Riemann B
 * fluxes and the face B fluxes from them (the same arithmetic as
 * ConstraintedTransport, so results are identical).  The flux arrays then
 * keep the Riemann B fluxes and the EMF arrays are not written by
 * pangu_stage.  0 runs ConstraintedTransport first (EMFs, then the face
 * fluxes), as the task-level entry points always do.  On a refined mesh it
 * always runs first, because flux correction overwrites some face fluxes. */
int pangu_set_fuse_ct(pangu_ctx *ctx, int fuse);
/* on = 1: with the exchange after the update (exchange_after 0) and
 * distinct input and output containers, pangu_stage runs the ghost
 * exchange of the output container on a second stream of the context from
 * the start of the stage, beside FixPrimitive, the flux sweeps and the
 * update, and the stage's stream waits for it before the recovery.  That
 * exchange reads the output container's interior P and flag and writes its
 * ghosts, which nothing before the recovery writes or reads, so results are
 * identical.  0 runs it on the stage's stream after the update.  -1
 * (default): 1 on a refined mesh, whose exchange is a large share of the
 * stage, 0 otherwise (on small meshes the cross-stream join costs more than
 * the overlap hides).  The task-level entry points are not affected. */
int pangu_set_overlap_exchange(pangu_ctx *ctx, int on);
/* on = 1 (default): pangu_stage runs the flux-divergence update (with the
 * geometric source) and the recovery (+ heating) of each cell in one
 * kernel, the recovery taking the new U from registers; the stage's
 * exchange (after the update) then runs before that kernel, which changes
 * nothing.  Applies with the chainer, Kastaun or 1Dvsq recovery (with a
 * learned start attached, its kernel runs the update first), no split
 * source and no persistent recovery, and on a magnetised mesh with the CT
 * face fluxes formed in the update (pangu_set_fuse_ct); the separate
 * kernels run otherwise.  Same operations per cell: results are identical.
 * 0: separate kernels. */
int pangu_set_fuse_update_recovery(pangu_ctx *ctx, int on);
/* on = 1: pangu_stage runs FixPrimitive + CalculateConservative and the X1
 * flux sweep in one kernel: a workgroup fixes a run of 248 cells of a
 * ghosted plane plus four halo cells on each side (fixed again, to the same
 * bits, by the neighbouring workgroup), keeps the fixed primitives in LDS
 * and forms the X1 faces of its cells from them.  Applies when the X1 sweep
 * uses its plain one-lane-per-face form.  Same operations: results are
 * identical.  0 (default): separate kernels. */
int pangu_set_fuse_fix_x1(pangu_ctx *ctx, int on);

/* ---- run diagnostics (no reference counterpart) ----
 * doc/grmhd_intro.md:556-574 defines the fluxes through a sphere that a
 * torus run is monitored by; the reference code computes none of them.
 * Evaluated on container md at the cell centres of one interior x1 index
 * per mesh (shell[m], 0 <= shell[m] < mesh cells along x1; -1 skips the
 * shell integrals, and a refined mesh accepts only -1, since its blocks
 * count x1 cells at their own levels), in code
 * coordinates over dx2 dx3 (a surface flux does not depend on them):
 *   MDOT = -sum rho u^1 sqrt|g|,  EDOT = sum T^1_0 sqrt|g|,
 *   JDOT = -sum T^1_3 sqrt|g|,    PHI  = 1/2 sum |B^1| sqrt|g|;
 * over the interior: MASS = sum U_rho dV, ENERGY = sum U_energy dV,
 * DIVB = max |corner divergence of U_B| (the quantity constrained transport
 * keeps at round-off), UBMAX = max |U_B| (its scale).  On a Cartesian
 * Kerr-Schild mesh the excised cells (r < r_excise, which Recovery skips and
 * whose U is not evolved) are left out of MASS, ENERGY and UBMAX, and
 * corners touching one out of DIVB.  out is a device array
 * of nmesh * PANGU_NDIAG doubles, written in stream order; the reductions run
 * in a fixed order, so repeated calls give identical bits.  shell is a host
 * array of one index per mesh. */
enum {
  PANGU_DIAG_MDOT = 0, PANGU_DIAG_EDOT, PANGU_DIAG_JDOT, PANGU_DIAG_PHI, PANGU_DIAG_MASS,
  PANGU_DIAG_ENERGY, PANGU_DIAG_DIVB, PANGU_DIAG_UBMAX, PANGU_NDIAG
};
int pangu_diagnostics(pangu_ctx *ctx, int md, const int *shell, double *out, void *stream);

/* pangu_diagnostics without any device allocation (no counterpart in the
 * reference).  For callers that cannot allocate device memory: code under
 * stream capture, tight serving loops, hosts that forbid hipMalloc.  Same
 * arguments, checks and results (bit for bit) as pangu_diagnostics, which
 * allocates its per-tile workspace on first use; this variant uses a static
 * workspace of 8 MiB per device instead (held until the process ends) and
 * returns PANGU_EINVAL when the context needs more: about 3.3e7 interior
 * zones (131 000 tiles of 256 cells, less one double per two meshes).  Every
 * call on a device shares that workspace, so calls that overlap on one
 * device (other host threads, other streams) race; calls on one stream from
 * one thread do not.  Synchronous, like pangu_diagnostics. */
int pangu_diagnostics_noalloc(pangu_ctx *ctx, int md, const int *shell, double *out, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* PANGU_HIP_H */
