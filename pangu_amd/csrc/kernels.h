// Launch interface between the C-ABI layer and the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/pangu_mlp.h"
#include "common.h"
#include "mlp_dev.h"
#include "physics_dev.h"

namespace pangu {

// Per-launch constants (passed by value; lives in kernel arguments).
struct KCtx {
  const DevBlock *blk;
  const DevGeo *geo;
  int nblocks, nmesh;
  int np, nc, iBX, iENT, iKEL;
  int ndim;            // largest dimensionality over the blocks
  int nplanes;         // interior (block, k) planes over all blocks
  const int *pblock;   // plane -> block
  const int *pfirst;   // first plane of each block (nblocks + 1 entries)
  // dt reduction items: each interior plane in chunks of 256 * dtper cells
  int ndt;
  int dtper;           // cells per thread of one item: 1, 4 or 16
  const int *dtblock;  // item -> block
  const int *dtfirst;  // first item of each block (nblocks + 1 entries)
  const int *mfirst;   // first block of each mesh (nmesh + 1 entries)
  Phys ph;
};

// Device pointer tables of one container: prim/cons [nblocks*ncomp],
// flux[d] [nblocks*ncons].
struct KFields {
  double *const *prim;
  double *const *cons;
  double *const *flux[3];
};

struct SharedTabs {
  double *const *flag;
  double *const *flag_rec;
  int32_t *const *status;
  int32_t *const *iters;
  double *const *root;  // Kastaun root of each recovered cell (NaN on failure), or null
  // learned Kastaun bracket (when the launch carries a model): relative
  // half-width of the first bracket, and the gate: a cell uses the model
  // when its previous recovery took at least `gate` false-position steps
  // (read from iters before it is overwritten; 0 = every cell)
  double delta;
  int gate;
};

// Launch range of a kernel over every block of a context, as 256-thread
// tiles.  Blocks may differ in shape: tblock maps a tile to its block and
// tfirst gives each block's first tile.  On an active dim d of a block the
// range is s[d] + rlo[d] .. e[d] + rhi[d] (entire: 0 .. n[d] - 1); an
// inactive dim has its single index.  A block's tiles cover its k planes,
// each plane flattened row-major.  order: 0 = plane-major (consecutive tiles
// sweep a plane), 1 = column-major (consecutive tiles walk k for one
// 256-cell run of the plane, reusing x3-independent geometry from L2).
struct Range3 {
  const int *tblock;
  const int *tfirst;
  int ntiles;
  int rlo[3], rhi[3];
  int entire;
  int order;
  // When every block of the context has one shape the tiles are decoded
  // arithmetically from these (no table loads): lo and cnt of the range on
  // each block, tiles per block plane.
  int uni, ulo[3], ucnt[3], utpp;
  // Marching ranges of a sweep (decode_march): faces per thread along the
  // sweep direction; 0 elsewhere.
  int step;
  // Cells per tile = threads per workgroup of the launch (256, 128 or 64).
  int tile;
};

// Work items counted per block (ghost cells, boundary slab cells), in tiles
// of 256 items: tile -> block and each block's first tile.
struct Items {
  const int *tblock;
  const int *tfirst;
  int ntiles;
};

// Optional per-launch clock: when a LaunchCtx carries one, every kernel
// launch is followed by a HIP event on the launch stream, so the duration of
// launch i is the time between events i and i+1.  Timing runs only: not for
// code under graph capture.
struct KernelClock {
  std::vector<const char *> names;  // names[i]: the launch ended by events[i + 1]
  std::vector<hipEvent_t> events;
  size_t n = 0;                     // events recorded since reset()
  hipError_t mark(const char *name, hipStream_t s);
  void reset() { n = 0; }
  ~KernelClock();
};

// Static mesh refinement tables (XFill, XProl, XFlux: common.h).
struct MultiLevel {
  int on = 0;
  const XFill *fill = nullptr;
  const XProl *prol = nullptr;
  const XFlux *flux = nullptr;
  int nfill = 0, nprol = 0, nflux = 0;
};

// Cells per thread of one dt reduction item for a mesh whose largest
// interior plane holds `plane` cells: an item (one workgroup of 256 threads)
// scans 256 * dt_per_thread(plane) cells of a plane, at most 4096 (a 64x64
// plane is one item, a 512x256 plane 32), with that many loads per thread
// in flight together; a mesh of small planes keeps the register count (and
// the workgroups resident) of one load per thread.
inline int dt_per_thread(int64_t plane) { return plane <= 256 ? 1 : (plane <= 1024 ? 4 : 16); }

// Host-side launch context: kernel constants plus the launch ranges.
struct LaunchCtx {
  KCtx k;
  int limiter, riemann, metric, recovery;
  Range3 interior, entire, face[3], ct;
  Range3 fixx1;  // the entire range in runs of 248 cells (k_fix_flux_x1)
  Range3 x1cells;                  // cell range of the shared-reconstruction X1 sweep
  Items ghosts, bcs[3];
  int order[PANGU_KF_COUNT] = {};  // tile order per kernel family (PANGU_ORDER_*)
  int split_source = 0;            // fused stage: source term in a kernel of its own
  int fuse_ct = 1;                 // fused stage: CT face fluxes formed in the update kernel
  int overlap_exchange = -1;       // fused stage: ghost exchange on the side stream (-1: refined meshes)
  int fuse_update_recovery = 1;    // fused stage: update and recovery in one kernel
  int fuse_fix_x1 = 0;             // fused stage: FixPrimitive + conservative and the X1 sweep in one kernel
  int share_x1 = 0;                // X1 sweep reconstructs each cell once (k_flux_x1s)
  int pair_flux = 0;               // bit d-1: the sweep along d runs two lanes per face (k_flux_pair)
  Range3 march[2];                 // marching ranges of the X2 and X3 sweeps (k_flux_march)
  // interior, entire, face[0..2] and ct again in tiles of 128 and 64 cells
  Range3 small[6][2];
  int tile[PANGU_KF_COUNT] = {};   // cells per tile of each kernel family (0: 256)
  int persist_recovery = 0;        // recovery as a persistent work-queue kernel (k_recover_queue)
  unsigned *queue = nullptr;       // its queue: next tile, workgroups done (zero between launches)
  int march_depth[2] = {};         // faces per thread of the X2, X3 sweeps; 0: one face per thread
  MultiLevel ml;                   // refinement tables (ml.on = 0 on a single-level mesh)
  KernelClock *clock = nullptr;
};

// Device view of a model's current weights in fragment order; false before
// any weights exist.
bool mlp_view(const pangu_mlp *m, MlpFrag &out);
// Appends the training pairs of container F (every stride-th interior cell,
// solved by the plain Kastaun recovery) to the model's data set; *kept
// receives the number of solved states added (pangu_mlp_harvest).
int mlp_harvest(pangu_mlp *m, const LaunchCtx &L, const KFields &F, int stride, int64_t *kept);
// Device the model lives on.
int mlp_device(const pangu_mlp *m);
// Target of the model: 0 Kastaun bracket, 1 1Dvsq start.
int mlp_kind(const pangu_mlp *m);

// Attaches (or, with nullptr, detaches) a clock to a context's launches.
void set_kernel_clock(pangu_ctx *c, KernelClock *clock);

int launch_fix_cons(const LaunchCtx &L, const KFields &F, uint8_t *const *mask, bool fix, bool cons,
                    hipStream_t s);
int launch_fluxes(const LaunchCtx &L, const KFields &F, double *const *alfven, hipStream_t s);
// Edge EMFs, then the CT face B fluxes (the task-level ConstraintedTransport).
int launch_ct(const LaunchCtx &L, const KFields &F, double *const *emf, hipStream_t s);
// mode: 1 divergence -> dudt, 2 update from dudt, 3 source only, 4 fused
// divergence + update + source, 5 fused divergence + update without source
// ct_fused (modes 4 and 5, with a field): the kernel forms the edge EMFs
// from the Riemann B fluxes and from them the constrained-transport face B
// fluxes, with the expressions of k_emf and k_ct_faces (neither is then
// launched); false reads the face B fluxes from the flux arrays.
int launch_div_update(const LaunchCtx &L, int mode, const KFields &F0, const KFields &Finit,
                      const KFields &F1, double *const *dudt, bool ct_fused,
                      const double *dt_dev, double beta, double bdt_host, hipStream_t s);
// rec: 0 chainer, 1 kastaun, 2 1dvsq, -1 none (heating only).  With a
// model W and rec 1, the recovery kernel evaluates the network in each
// wavefront whose cells pass the gate and starts those cells' Kastaun
// search from the learned bracket.
// Inputs of the mode-4 update a learned recovery kernel runs first when the
// stage fuses the two (see k_update_recover).
struct UpdArgs {
  KFields F0, Finit;
  const double *dt_dev = nullptr;
  double beta = 0.0, bdt_host = 0.0;
};
// upd (optional): the learned kernels run the cell's update first (the
// plain schemes then take launch_update_recover).
int launch_recover(const LaunchCtx &L, const KFields &F, const SharedTabs &S, int rec, bool heat,
                   const MlpFrag *W, hipStream_t s, const UpdArgs *upd = nullptr);
// The mode-4 update and the recovery (+ heating) of each cell in one kernel
// (k_update_recover): the chainer, Kastaun or 1Dvsq recovery, with the CT
// face fluxes formed in the update on a magnetised mesh, and K_e only with
// a field (-1 otherwise).
// FixPrimitive + CalculateConservative and the X1 sweep in one kernel
// (k_fix_flux_x1), and the X2 and X3 sweeps alone.
int launch_fix_flux_x1(const LaunchCtx &L, const KFields &F, uint8_t *const *mask, double *const *alfven,
                       hipStream_t s);
int launch_fluxes_23(const LaunchCtx &L, const KFields &F, double *const *alfven, hipStream_t s);
int launch_update_recover(const LaunchCtx &L, const KFields &F0, const KFields &Finit,
                          const KFields &F1, const SharedTabs &S, int rec, bool heat, bool ct_fused,
                          const double *dt_dev, double beta, double bdt_host, hipStream_t s);
// Training pairs from a container: U in the recovery's normalisation and
// the cell's metric sample (35 doubles) of every stride-th interior cell
// (stride >= 1), solved by the plain Kastaun recovery (kind 1: then by
// 1Dvsq from Kastaun's v^2, with entropy U_ent / D).  Writes slot
// t = (tile * 256 + thread) / stride for t < cap: U [t][9], mtab [t][35],
// feats [t][5], target [t] (logit of the root; kind 1: log u^2), iters [t],
// status [t] (0 solved, < 0 no cell / excised / unsolved, 10 + code a
// failed 1Dvsq).  Returns the slots written.
int64_t launch_harvest(const LaunchCtx &L, const KFields &F, int stride, double gamma, int kind,
                       int64_t cap, double *U, double *mtab, float *feats, float *target, int *iters,
                       int *status, hipStream_t s);
int launch_fix_recovery(const LaunchCtx &L, const KFields &F, double *const *flag, int *cluster,
                        hipStream_t s);
int launch_exchange(const LaunchCtx &L, const KFields &F, double *const *flag, const int *nbr,
                    hipStream_t s);
int launch_bc(const LaunchCtx &L, const KFields &F, double *const *flag, hipStream_t s);
// Refined mesh: ghost cells across and within levels (fills, then
// prolongation), and coarse face fluxes from the fine ones.
int launch_exchange_ml(const LaunchCtx &L, const KFields &F, double *const *flag, hipStream_t s);
int launch_flux_correction(const LaunchCtx &L, const KFields &F, hipStream_t s);
// red: 3 doubles of scratch per interior plane; state: {dt, time} per mesh;
// est_out (optional): one estimate per mesh
int launch_diagnostics(const LaunchCtx &L, const KFields &F, const int *shell, double *part,
                       const int *mtile, int nmesh, double *out, hipStream_t s);
int launch_dt(const LaunchCtx &L, double *const *alfven, double *red, double *state,
              double dt_factor, double *est_out, hipStream_t s);

// Static workspace of the allocation-free diagnostics entry point (diag.hip):
// 2^20 doubles (8 MiB) per device, enough for the per-tile partials of
// 131 000 interior tiles (about 3.3e7 interior zones) plus the per-mesh ints.
constexpr int64_t kDiagScratchDoubles = int64_t(1) << 20;
constexpr int kDiagMaxDevices = 64;
// Device address of that workspace on `device`, which must be the current
// device; looked up once per device (hipGetSymbolAddress) and kept.
hipError_t diag_scratch_address(int device, double **out);

}  // namespace pangu
