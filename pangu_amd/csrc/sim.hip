// Host driver above the pangu_hip C ABI (see include/pangu_sim.h).
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/pangu_hip.h"
#include "../../include/pangu_sim.h"
#include "host.h"
#include "kernels.h"

using namespace pangu;

// field ids (shared with the test harness)
enum {
  F_PRIM = 0, F_CONS = 1, F_FLUX1 = 2, F_FLUX2 = 3, F_FLUX3 = 4, F_ALFVEN = 5, F_EMF = 6,
  F_FLAG = 7, F_MASK = 8, F_STATUS = 9, F_ITERS = 10, F_FLAG_REC = 11, F_DUDT = 12, F_ROOT = 13
};
// task ids
enum {
  T_FIX = 0, T_CONS, T_FLUXES, T_CT, T_DIV, T_UPDATE, T_SOURCE, T_RECOVERY, T_HEATING,
  T_FIXREC, T_EXCHANGE, T_BC, T_TIMESTEP, T_FLUXCORR
};

struct pangu_sim {
  pangu_sim_cfg cfg;
  int device = 0;
  pangu_ctx *ctx = nullptr;
  std::vector<pangu_block> blocks;
  std::vector<HostBlock> hb;
  int nblk = 0, np = 0, nc = 0, iBX = -1, iENT = 0, iKEL = -1;
  int64_t ncell = 0, total = 0;
  double *P[2] = {nullptr, nullptr}, *U[2] = {nullptr, nullptr}, *dudt = nullptr;
  double *F[3] = {nullptr, nullptr, nullptr};
  double *alfven = nullptr, *emf = nullptr, *flag = nullptr, *flag_rec = nullptr;
  uint8_t *mask = nullptr;
  int32_t *status = nullptr, *iters = nullptr;
  double *root = nullptr;    // Kastaun roots, allocated while recording (pangu_sim_record_roots)
  int h_base = -1, h_s1 = -1, h_dudt = -1;
  double *dstate = nullptr;  // {dt, time} per mesh
  double *ddiag = nullptr;   // diagnostics output, PANGU_NDIAG per mesh
  double dt = 0, time = 0, dt_est = DBL_MAX;  // mesh 0 (task mode: the only mesh)
  int nmesh = 1;
  std::vector<pangu_sim_cfg> mcfg;  // per mesh
  std::vector<int> bmesh;           // mesh of each block
  std::vector<int64_t> boff;        // first cell of each block in the arenas
  std::vector<double> mdt, mtime;   // per mesh dt and time (host copy)
  int ncycle = 0;
  bool stage1_ready = false;
  bool dt_on_device = false;  // the fused path owns dt/time on the device
  hipStream_t stream = nullptr;
  hipGraphExec_t gexec = nullptr;
  hipGraph_t graph = nullptr;
  std::vector<std::string> knames;  // kernels of the last pangu_sim_time_kernels
  std::string err;
};

static int sfail(pangu_sim *s, int code, const std::string &m) {
  s->err = m;
  return code;
}
#define SCHK(s, call)                                                                       \
  do {                                                                                      \
    hipError_t e_ = (call);                                                                 \
    if (e_ != hipSuccess) return sfail(s, PANGU_EDEVICE, std::string(#call ": ") + hipGetErrorString(e_)); \
  } while (0)
#define PCHK(s, call)                                                               \
  do {                                                                              \
    int r_ = (call);                                                                \
    if (r_ != 0) return sfail(s, r_, std::string(#call ": ") + pangu_last_error(s->ctx)); \
  } while (0)

// Parthenon/Athena++ uniform mesh generator for interior block edges; mesh
// edges are exact.
static double mesh_gen(int idx, int n, double xmin, double xmax) {
  const long noff = idx - n / 2;
  const long noffc = idx - (n + 1) / 2;
  const double x = (double)(noff + noffc) / (2.0 * n);
  return 0.5 * (xmin + xmax) + (x * xmax - x * xmin);
}

// Edge idx of n equal cells over [xmin, xmax] (mesh edges exact).
static double edge_at(int64_t idx, int64_t n, double xmin, double xmax) {
  return idx == 0 ? xmin : (idx == n ? xmax : mesh_gen((int)idx, (int)n, xmin, xmax));
}

// Block tree of a statically refined mesh (Parthenon's Mesh constructor with
// <parthenon/static_refinementN> blocks).  Nodes are (level, lx); a refined
// node has 2^dim children ordered x1 fastest.  Refining a node first makes
// every neighbour (faces, edges, corners; periodic wrap) at least as fine as
// the node, so leaves next to each other differ by at most one level.
struct LeafTree {
  struct Node {
    int level;
    int64_t l[3];
    int child0;  // first child, -1 for a leaf
  };
  std::vector<Node> nodes;  // root nodes first, x1 fastest
  int dim = 1;
  int64_t nb[3] = {1, 1, 1};
  bool periodic[3] = {false, false, false};

  LeafTree(const pangu_sim_cfg *c) {
    dim = c->nx[2] > 1 ? 3 : (c->nx[1] > 1 ? 2 : 1);
    for (int d = 0; d < 3; ++d) {
      nb[d] = c->nb[d];
      periodic[d] = c->bc[d][0] == PANGU_BC_PERIODIC;
    }
    for (int64_t k = 0; k < nb[2]; ++k)
      for (int64_t j = 0; j < nb[1]; ++j)
        for (int64_t i = 0; i < nb[0]; ++i) nodes.push_back({0, {i, j, k}, -1});
  }
  // Deepest node on the path to (level, l): the node itself or the leaf
  // that covers it.
  int descend(int level, const int64_t l[3]) const {
    int64_t r[3];
    for (int d = 0; d < 3; ++d) r[d] = d < dim ? l[d] >> level : 0;
    int n = (int)((r[2] * nb[1] + r[1]) * nb[0] + r[0]);
    while (nodes[n].level < level && nodes[n].child0 >= 0) {
      const int sh = level - nodes[n].level - 1;
      int c = 0;
      for (int d = 0; d < dim; ++d) c |= (int)((l[d] >> sh) & 1) << d;
      n = nodes[n].child0 + c;
    }
    return n;
  }
  void refine(int n) {
    const int L = nodes[n].level;
    // neighbours of n must reach level L first
    for (int o2 = (dim > 2 ? -1 : 0); o2 <= (dim > 2 ? 1 : 0); ++o2)
      for (int o1 = (dim > 1 ? -1 : 0); o1 <= (dim > 1 ? 1 : 0); ++o1)
        for (int o0 = -1; o0 <= 1; ++o0) {
          const int o[3] = {o0, o1, o2};
          int64_t l[3];
          bool inside = true;
          for (int d = 0; d < 3; ++d) {
            l[d] = nodes[n].l[d] + o[d];
            const int64_t N = d < dim ? nb[d] << L : 1;
            if (l[d] < 0 || l[d] >= N) {
              if (periodic[d] && d < dim)
                l[d] = (l[d] + N) % N;
              else
                inside = false;
            }
          }
          if (inside && nodes[descend(L, l)].level < L) add(L, l);
        }
    if (nodes[n].child0 >= 0) return;  // refined meanwhile through a neighbour
    const int first = (int)nodes.size();
    for (int c = 0; c < (1 << dim); ++c) {
      Node ch;
      ch.level = L + 1;
      ch.child0 = -1;
      for (int d = 0; d < 3; ++d) ch.l[d] = d < dim ? 2 * nodes[n].l[d] + ((c >> d) & 1) : 0;
      nodes.push_back(ch);
    }
    nodes[n].child0 = first;
  }
  // Makes (level, l) a node of the tree.
  void add(int level, const int64_t l[3]) {
    for (;;) {
      const int n = descend(level, l);
      if (nodes[n].level == level) return;
      refine(n);
    }
  }
  void leaves(int n, std::vector<int> &out) const {
    if (nodes[n].child0 < 0) {
      out.push_back(n);
      return;
    }
    for (int c = 0; c < (1 << dim); ++c) leaves(nodes[n].child0 + c, out);
  }
};

// Leaves of a mesh after its refinement regions (in the order given).  Per
// active dim at level L (N = nb << L blocks), a region spans the blocks from
// the first whose upper edge lies above its lower bound to the first whose
// upper edge reaches its upper bound, widened to whole sibling pairs.
static int mesh_leaves(const pangu_sim_cfg *c, std::vector<pangu_block> &out, std::string &err) {
  if (c->nref < 0 || c->nref > PANGU_SIM_MAX_REF) {
    err = "nref out of range";
    return PANGU_EINVAL;
  }
  LeafTree T(c);
  for (int r = 0; r < c->nref; ++r) {
    const int L = c->ref_level[r];
    if (L < 0 || L > 20) {
      err = "refinement level out of range (0..20)";
      return PANGU_EINVAL;
    }
    if (L == 0) continue;
    int64_t lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    bool empty = false;
    for (int d = 0; d < T.dim; ++d) {
      const int64_t N = T.nb[d] << L;
      int64_t a = 0;
      while (a < N && !(edge_at(a + 1, N, c->xmin[d], c->xmax[d]) > c->ref_min[r][d])) ++a;
      int64_t b = a;
      while (b < N && !(edge_at(b + 1, N, c->xmin[d], c->xmax[d]) >= c->ref_max[r][d])) ++b;
      if (a >= N) empty = true;
      if (b >= N) b = N - 1;
      a -= a & 1;
      if (!(b & 1)) ++b;
      lo[d] = a;
      hi[d] = b;
    }
    if (empty) continue;
    for (int64_t k = lo[2]; k <= hi[2]; k += (T.dim > 2 ? 2 : 1))
      for (int64_t j = lo[1]; j <= hi[1]; j += (T.dim > 1 ? 2 : 1))
        for (int64_t i = lo[0]; i <= hi[0]; i += 2) {
          const int64_t l[3] = {i, j, k};
          T.add(L, l);
        }
  }
  std::vector<int> lv;
  const int nroot = (int)(T.nb[0] * T.nb[1] * T.nb[2]);
  for (int n = 0; n < nroot; ++n) T.leaves(n, lv);
  out.clear();
  for (int n : lv) {
    const LeafTree::Node &N = T.nodes[n];
    pangu_block B;
    std::memset(&B, 0, sizeof(B));
    B.level = N.level;
    for (int d = 0; d < 3; ++d) {
      B.lx[d] = (int)N.l[d];
      B.nx[d] = c->nx[d] / c->nb[d];
      B.ng[d] = (c->nx[d] > 1) ? c->ng : 0;
      const int64_t cnt = d < T.dim ? T.nb[d] << N.level : 1;
      const double lo = edge_at(N.l[d], cnt, c->xmin[d], c->xmax[d]);
      const double hi = edge_at(N.l[d] + 1, cnt, c->xmin[d], c->xmax[d]);
      B.xmin[d] = lo;
      B.dx[d] = (hi - lo) / B.nx[d];
    }
    out.push_back(B);
  }
  return 0;
}

template <class T>
static int dalloc(pangu_sim *s, T **p, int64_t n) {
  SCHK(s, hipMalloc((void **)p, (size_t)(n > 0 ? n : 1) * sizeof(T)));
  SCHK(s, hipMemset(*p, 0, (size_t)(n > 0 ? n : 1) * sizeof(T)));
  return 0;
}

static void sync_dt_from_device(pangu_sim *s) {
  if (!s->dt_on_device) return;
  std::vector<double> st(2 * (size_t)s->nmesh);
  (void)hipMemcpy(st.data(), s->dstate, st.size() * sizeof(double), hipMemcpyDeviceToHost);
  for (int m = 0; m < s->nmesh; ++m) {
    s->mdt[m] = st[2 * m];
    s->mtime[m] = st[2 * m + 1];
  }
  s->dt = s->mdt[0];
  s->time = s->mtime[0];
}

static void push_dt_to_device(pangu_sim *s) {
  s->mdt[0] = s->dt;
  s->mtime[0] = s->time;
  std::vector<double> st(2 * (size_t)s->nmesh);
  for (int m = 0; m < s->nmesh; ++m) {
    st[2 * m] = s->mdt[m];
    st[2 * m + 1] = s->mtime[m];
  }
  (void)hipMemcpy(s->dstate, st.data(), st.size() * sizeof(double), hipMemcpyHostToDevice);
}

static pangu_config physics_of(const pangu_sim_cfg *c) {
  pangu_config pc;
  std::memset(&pc, 0, sizeof(pc));
  pc.metric = c->metric;
  pc.a = c->a;
  pc.h = c->h;
  pc.gamma = c->gamma;
  pc.cfl = c->cfl;
  pc.riemann = c->riemann;
  pc.limiter = c->limiter;
  pc.recovery = c->recovery;
  pc.enable_B = c->enable_B;
  pc.enable_heating = c->enable_heating;
  pc.density_floor = c->density_floor;
  pc.energy_floor = c->energy_floor;
  pc.density_floor_pow = c->density_floor_pow;
  pc.energy_floor_pow = c->energy_floor_pow;
  pc.sigma_max = c->sigma_max;
  pc.lorentz_max = c->lorentz_max;
  pc.heat_model = c->heat_model;
  pc.fel_constant = c->fel_constant;
  pc.gamma_e = c->gamma_e;
  pc.gamma_p = c->gamma_p;
  pc.ratio_min = c->ratio_min;
  pc.ratio_max = c->ratio_max;
  pc.limit_kel = c->limit_kel;
  pc.suppress_highb_heat = c->suppress_highb_heat;
  pc.enforce_positive_dissipation = c->enforce_positive_dissipation;
  pc.excise = c->excise;
  pc.r_excise = c->r_excise;
  pc.dexcise = c->dexcise;
  pc.pexcise = c->pexcise;
  for (int d = 0; d < 3; ++d) {
    pc.nb[d] = c->nb[d];
    pc.bc[d][0] = c->bc[d][0];
    pc.bc[d][1] = c->bc[d][1];
  }
  return pc;
}

extern "C" {

// Registers the shared fields (wave speeds, EMFs, flags, masks, recovery
// codes and, while recording, the Kastaun roots) with the context.
static int register_shared(pangu_sim *s) {
  auto table = [&](double *base, int ncomp) {
    std::vector<double *> t((size_t)s->nblk * ncomp);
    for (int b = 0; b < s->nblk; ++b)
      for (int q = 0; q < ncomp; ++q)
        t[(size_t)b * ncomp + q] = base + s->boff[b] * ncomp + (int64_t)q * s->hb[b].cells;
    return t;
  };
  std::vector<double *> ta = table(s->alfven, 3), te = table(s->emf, 3), tfl = table(s->flag, 1),
                        tfr = table(s->flag_rec, 1);
  std::vector<double *> tr = s->root ? table(s->root, 1) : std::vector<double *>();
  std::vector<uint8_t *> tm(s->nblk);
  std::vector<int32_t *> tst(s->nblk), tit(s->nblk);
  for (int b = 0; b < s->nblk; ++b) {
    tm[b] = s->mask + s->boff[b];
    tst[b] = s->status + s->boff[b];
    tit[b] = s->iters + s->boff[b];
  }
  pangu_shared sh;
  sh.alfven = ta.data();
  sh.emf = te.data();
  sh.flag = tfl.data();
  sh.mask = tm.data();
  sh.c2p_status = tst.data();
  sh.c2p_iters = tit.data();
  sh.flag_recovery = tfr.data();
  sh.c2p_root = s->root ? tr.data() : nullptr;
  return pangu_register_shared(s->ctx, &sh);
}

int pangu_sim_create_batch(const pangu_sim_cfg *cfgs, int nmesh, int device, pangu_sim **out) {
  if (!cfgs || nmesh <= 0 || !out) return PANGU_EINVAL;
  *out = nullptr;
  for (int m = 0; m < nmesh; ++m) {
    const pangu_sim_cfg *c = &cfgs[m];
    for (int d = 0; d < 3; ++d)
      if (c->nb[d] <= 0 || c->nx[d] % c->nb[d] != 0) return PANGU_EINVAL;
    if (c->dt_factor != cfgs[0].dt_factor || c->stage_init != cfgs[0].stage_init ||
        c->exchange_after != cfgs[0].exchange_after) {
      fprintf(stderr, "pangu_sim_create_batch: meshes must share dt_factor, stage_init, exchange_after\n");
      return PANGU_EINVAL;
    }
  }
  const pangu_sim_cfg *c = &cfgs[0];
  pangu_sim *s = new pangu_sim();
  s->cfg = *c;
  s->mcfg.assign(cfgs, cfgs + nmesh);
  s->nmesh = nmesh;
  s->device = device;
  if (hipSetDevice(device) != hipSuccess) {
    delete s;
    return PANGU_EDEVICE;
  }
  for (int m = 0; m < nmesh; ++m) {
    const pangu_sim_cfg *cm = &cfgs[m];
    std::vector<pangu_block> leaves;
    if (int r = mesh_leaves(cm, leaves, s->err)) {
      fprintf(stderr, "pangu_sim_create: %s\n", s->err.c_str());
      delete s;
      return r;
    }
    for (const pangu_block &B : leaves) {
      s->blocks.push_back(B);
      s->hb.push_back(make_host_block(B));
      s->bmesh.push_back(m);
    }
    s->mdt.push_back(cm->dt_init);
    s->mtime.push_back(0.0);
  }
  s->nblk = (int)s->blocks.size();
  std::vector<pangu_config> pcs(nmesh);
  for (int m = 0; m < nmesh; ++m) pcs[m] = physics_of(&cfgs[m]);
  int r = pangu_create_batch(pcs.data(), nmesh, s->blocks.data(), s->bmesh.data(), s->nblk, device,
                             &s->ctx);
  if (r) {
    delete s;
    return r;
  }
  s->np = 5;
  if (c->enable_B) {
    s->iBX = s->np;
    s->np += 3;
  }
  s->iENT = s->np++;
  if (c->enable_heating) s->iKEL = s->np++;
  s->nc = s->np;
  s->ncell = s->hb[0].cells;
  s->total = 0;
  for (int b = 0; b < s->nblk; ++b) {
    s->boff.push_back(s->total);
    s->total += s->hb[b].cells;
  }
  const int64_t T = s->total;
  if (dalloc(s, &s->P[0], T * s->np) || dalloc(s, &s->P[1], T * s->np) ||
      dalloc(s, &s->U[0], T * s->nc) || dalloc(s, &s->U[1], T * s->nc) ||
      dalloc(s, &s->dudt, T * s->nc) || dalloc(s, &s->F[0], T * s->nc) ||
      dalloc(s, &s->F[1], T * s->nc) || dalloc(s, &s->F[2], T * s->nc) ||
      dalloc(s, &s->alfven, T * 3) || dalloc(s, &s->emf, T * 3) || dalloc(s, &s->flag, T) ||
      dalloc(s, &s->flag_rec, T) || dalloc(s, &s->mask, T) || dalloc(s, &s->status, T) ||
      dalloc(s, &s->iters, T) || dalloc(s, &s->dstate, 2 * (int64_t)nmesh)) {
    fprintf(stderr, "pangu_sim_create: %s\n", s->err.c_str());
    pangu_sim_destroy(s);
    return PANGU_ENOMEM;
  }
  auto table = [&](double *base, int ncomp) {
    std::vector<double *> t((size_t)s->nblk * ncomp);
    for (int b = 0; b < s->nblk; ++b)
      for (int q = 0; q < ncomp; ++q)
        t[(size_t)b * ncomp + q] = base + s->boff[b] * ncomp + (int64_t)q * s->hb[b].cells;
    return t;
  };
  auto register_cont = [&](double *Pa, double *Ua, bool with_flux, int *h) {
    std::vector<double *> tp = table(Pa, s->np), tu = table(Ua, s->nc);
    std::vector<double *> tf[3];
    pangu_container cc;
    cc.prim = tp.data();
    cc.cons = tu.data();
    for (int d = 0; d < 3; ++d) {
      tf[d] = table(s->F[d], s->nc);
      cc.flux[d] = with_flux ? tf[d].data() : nullptr;
    }
    return pangu_register_container(s->ctx, &cc, h);
  };
  // Fluxes are consumed within the stage that computes them, so the two
  // stage containers share one set of flux arrays.
  r = register_cont(s->P[0], s->U[0], true, &s->h_base);
  if (!r) r = register_cont(s->P[1], s->U[1], true, &s->h_s1);
  if (!r) r = register_cont(s->P[1], s->dudt, false, &s->h_dudt);
  if (!r) r = register_shared(s);
  if (r) {
    fprintf(stderr, "pangu_sim_create: %s\n", pangu_last_error(s->ctx));
    pangu_sim_destroy(s);
    return r;
  }
  if (hipStreamCreate(&s->stream) != hipSuccess) {
    pangu_sim_destroy(s);
    return PANGU_EDEVICE;
  }
  s->dt = c->dt_init;
  s->time = 0.0;
  push_dt_to_device(s);
  *out = s;
  return PANGU_OK;
}

int pangu_sim_create(const pangu_sim_cfg *c, int device, pangu_sim **out) {
  return pangu_sim_create_batch(c, 1, device, out);
}

// Drops the captured cycle graph (after a change of launch parameters).
static void drop_graph(pangu_sim *s) {
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  if (s->gexec) (void)hipGraphExecDestroy(s->gexec);
  if (s->graph) (void)hipGraphDestroy(s->graph);
  s->gexec = nullptr;
  s->graph = nullptr;
}

int pangu_sim_destroy(pangu_sim *s) {
  if (!s) return 0;
  (void)hipSetDevice(s->device);
  drop_graph(s);
  void *ptrs[] = {s->P[0], s->P[1], s->U[0], s->U[1], s->dudt, s->F[0], s->F[1], s->F[2],
                  s->alfven, s->emf, s->flag, s->flag_rec, s->mask, s->status, s->iters, s->dstate,
                  s->ddiag};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  if (s->ctx) pangu_destroy(s->ctx);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
  return 0;
}

int pangu_sim_init(pangu_sim *s) {
  int first = 0;
  for (int m = 0; m < s->nmesh; ++m) {
    const pangu_sim_cfg &cm = s->mcfg[m];
    int last = first;
    while (last < s->nblk && s->bmesh[last] == m) ++last;
    ProblemParams pp;
    pp.problem = cm.problem;
    pp.rin = cm.rin;
    pp.rmax = cm.rmax;
    pp.kappa = cm.kappa;
    pp.perturbation = cm.perturbation;
    pp.beta = cm.beta;
    pp.aphi_cut = cm.aphi_cut;
    pp.fel_0 = cm.fel_0;
    pp.seed = cm.seed;
    pp.mad = cm.magnetic_mad;
    for (int d = 0; d < 3; ++d) pp.mesh_nx[d] = cm.nx[d];
    const pangu_config pc = physics_of(&cm);
    const std::vector<HostBlock> hb(s->hb.begin() + first, s->hb.begin() + last);
    std::vector<double> prim;
    if (generate_problem(pc, pp, hb, s->np, s->iBX, s->iENT, s->iKEL, prim))
      return sfail(s, PANGU_EINVAL, "unknown problem");
    SCHK(s, hipMemcpy(s->P[0] + s->boff[first] * s->np, prim.data(), prim.size() * sizeof(double),
                      hipMemcpyHostToDevice));
    first = last;
  }
  PCHK(s, pangu_exchange_ghosts(s->ctx, s->h_base, s->stream));
  PCHK(s, pangu_apply_boundary_conditions(s->ctx, s->h_base, s->stream));
  SCHK(s, hipStreamSynchronize(s->stream));
  return 0;
}

static int ensure_stage1(pangu_sim *s) {
  if (s->stage1_ready) return 0;
  if (s->cfg.stage_init == 1)
    SCHK(s, hipMemcpyAsync(s->P[1], s->P[0], (size_t)(s->total * s->np) * sizeof(double),
                           hipMemcpyDeviceToDevice, s->stream));
  s->stage1_ready = true;
  return 0;
}

// One RK stage at the reference's task granularity (ideal_grmhd.cc:52-116).
static int stage_tasks(pangu_sim *s, int stage) {
  const double beta = (stage == 1) ? 0.5 : 1.0;
  const int c0 = (stage == 1) ? s->h_base : s->h_s1;
  const int c1 = (stage == 1) ? s->h_s1 : s->h_base;
  const double bdt = beta * s->dt;
  pangu_ctx *x = s->ctx;
  void *st = s->stream;
  PCHK(s, pangu_fix_primitive(x, c0, st));
  PCHK(s, pangu_calculate_conservative(x, c0, st));
  PCHK(s, pangu_calculate_fluxes(x, c0, st));
  if (s->cfg.enable_B) PCHK(s, pangu_constrained_transport(x, c0, st));
  PCHK(s, pangu_flux_correction(x, c0, st));
  PCHK(s, pangu_flux_divergence(x, c0, s->h_dudt, st));
  PCHK(s, pangu_update_independent(x, s->h_base, s->h_dudt, bdt, c1, st));
  if (s->cfg.exchange_after == 0) PCHK(s, pangu_exchange_ghosts(x, c1, st));
  PCHK(s, pangu_add_geometric_source(x, c1, bdt, st));
  PCHK(s, pangu_recovery(x, c1, st));
  if (s->cfg.enable_heating) PCHK(s, pangu_apply_electron_heating(x, c1, st));
  if (s->cfg.recovery != PANGU_RECOVERY_ROBUSTER) PCHK(s, pangu_fix_recovery(x, c1, st));
  if (s->cfg.exchange_after == 1) PCHK(s, pangu_exchange_ghosts(x, c1, st));
  PCHK(s, pangu_apply_boundary_conditions(x, c1, st));
  if (stage == 2) {
    std::vector<double> est(s->nmesh);
    PCHK(s, pangu_estimate_timestep(x, c1, est.data(), st));
    s->dt_est = est[0];
  }
  return 0;
}

static int cycle_fused(pangu_sim *s) {
  PCHK(s, pangu_stage(s->ctx, s->h_base, s->h_base, s->h_s1, 0.5, s->dstate, s->cfg.dt_factor, 0,
                      s->cfg.exchange_after, s->stream));
  PCHK(s, pangu_stage(s->ctx, s->h_base, s->h_s1, s->h_base, 1.0, s->dstate, s->cfg.dt_factor, 1,
                      s->cfg.exchange_after, s->stream));
  return 0;
}

// Captures one fused cycle as a graph (once per set of launch parameters).
static int ensure_graph(pangu_sim *s) {
  if (s->gexec) return 0;
  SCHK(s, hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
  const int r = cycle_fused(s);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(s->stream, &g);
  if (r) return r;
  if (e != hipSuccess) return sfail(s, PANGU_EDEVICE, "hipStreamEndCapture failed");
  s->graph = g;
  SCHK(s, hipGraphInstantiate(&s->gexec, g, nullptr, nullptr, 0));
  return 0;
}

int pangu_sim_step(pangu_sim *s, int n, int mode) {
  SCHK(s, hipSetDevice(s->device));
  if (int r = ensure_stage1(s)) return r;
  if (mode == PANGU_SIM_TASKS) {
    if (s->nmesh != 1) return sfail(s, PANGU_EINVAL, "task mode runs one mesh (host dt)");
    if (s->dt_on_device) {
      sync_dt_from_device(s);
      s->dt_on_device = false;
    }
    for (int c = 0; c < n; ++c) {
      if (int r = stage_tasks(s, 1)) return r;
      if (int r = stage_tasks(s, 2)) return r;
      s->time += s->dt;
      const double grown = s->cfg.dt_factor * s->dt;
      s->dt = (s->dt_est < grown) ? s->dt_est : grown;
      s->ncycle++;
    }
    return 0;
  }
  if (!s->dt_on_device) {
    push_dt_to_device(s);
    s->dt_on_device = true;
  }
  if (mode == PANGU_SIM_GRAPH) {
    if (int r = ensure_graph(s)) return r;
    for (int c = 0; c < n; ++c) SCHK(s, hipGraphLaunch(s->gexec, s->stream));
  } else {
    for (int c = 0; c < n; ++c)
      if (int r = cycle_fused(s)) return r;
  }
  s->ncycle += n;
  return 0;
}

int pangu_sim_sync(pangu_sim *s) {
  SCHK(s, hipStreamSynchronize(s->stream));
  return 0;
}

int pangu_sim_timed_cycles(pangu_sim *s, int n, int mode, double *ms) {
  hipEvent_t a, b;
  SCHK(s, hipEventCreate(&a));
  SCHK(s, hipEventCreate(&b));
  SCHK(s, hipEventRecord(a, s->stream));
  if (int r = pangu_sim_step(s, n, mode)) return r;
  SCHK(s, hipEventRecord(b, s->stream));
  SCHK(s, hipEventSynchronize(b));
  float f = 0.f;
  SCHK(s, hipEventElapsedTime(&f, a, b));
  *ms = f;
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return 0;
}

// Runs ncycles fused cycles with the kernel clock: per kernel name the total
// milliseconds and launches over all cycles, in order of first launch.
static int clock_cycles(pangu_sim *s, int ncycles, std::vector<std::string> &names,
                        std::vector<double> &tot, std::vector<int> &calls) {
  if (int r = ensure_stage1(s)) return r;
  if (!s->dt_on_device) {
    push_dt_to_device(s);
    s->dt_on_device = true;
  }
  KernelClock clk;
  for (int c = 0; c < ncycles; ++c) {
    clk.reset();
    SCHK(s, clk.mark(nullptr, s->stream));
    set_kernel_clock(s->ctx, &clk);
    const int r = cycle_fused(s);
    set_kernel_clock(s->ctx, nullptr);
    if (r) return r;
    SCHK(s, hipStreamSynchronize(s->stream));
    for (size_t i = 1; i < clk.n; ++i) {
      float f = 0.f;
      SCHK(s, hipEventElapsedTime(&f, clk.events[i - 1], clk.events[i]));
      const std::string nm = clk.names[i - 1] ? clk.names[i - 1] : "?";
      size_t j = 0;
      while (j < names.size() && names[j] != nm) ++j;
      if (j == names.size()) {
        names.push_back(nm);
        tot.push_back(0.0);
        calls.push_back(0);
      }
      tot[j] += f;
      calls[j] += 1;
    }
  }
  s->ncycle += ncycles;
  return 0;
}

int pangu_sim_time_kernels(pangu_sim *s, int ncycles, double *ms_out, int *calls_out, int maxk) {
  (void)hipSetDevice(s->device);
  if (ncycles < 1 || !ms_out || !calls_out) return sfail(s, PANGU_EINVAL, "bad arguments");
  std::vector<std::string> names;
  std::vector<double> tot;
  std::vector<int> calls;
  if (int r = clock_cycles(s, ncycles, names, tot, calls)) return r;
  s->knames = names;
  const int nk = (int)names.size();
  for (int j = 0; j < nk && j < maxk; ++j) {
    ms_out[j] = tot[j] / ncycles;
    calls_out[j] = calls[j] / ncycles;
  }
  return nk;
}

int pangu_sim_plain_loop(pangu_sim **sims, int n, int ncycles, int mode, double *ms) {
  if (!sims || n <= 0 || ncycles < 0 || !ms || (mode != PANGU_SIM_FUSED && mode != PANGU_SIM_GRAPH))
    return PANGU_EINVAL;
  pangu_sim *s0 = sims[0];
  (void)hipSetDevice(s0->device);
  for (int i = 0; i < n; ++i) {
    pangu_sim *s = sims[i];
    if (s->device != s0->device) return sfail(s0, PANGU_EINVAL, "simulations on different devices");
    if (int r = ensure_stage1(s)) return r;
    if (!s->dt_on_device) {
      push_dt_to_device(s);
      s->dt_on_device = true;
    }
    if (mode == PANGU_SIM_GRAPH)
      if (int r = ensure_graph(s)) return r;
    SCHK(s, hipStreamSynchronize(s->stream));
  }
  hipStream_t st = s0->stream;
  hipEvent_t a, b;
  SCHK(s0, hipEventCreate(&a));
  SCHK(s0, hipEventCreate(&b));
  SCHK(s0, hipEventRecord(a, st));
  for (int c = 0; c < ncycles; ++c)
    for (int i = 0; i < n; ++i) {
      pangu_sim *s = sims[i];
      if (mode == PANGU_SIM_GRAPH) {
        SCHK(s0, hipGraphLaunch(s->gexec, st));
      } else {
        hipStream_t own = s->stream;
        s->stream = st;
        const int r = cycle_fused(s);
        s->stream = own;
        if (r) return r;
      }
      s->ncycle++;
    }
  SCHK(s0, hipEventRecord(b, st));
  SCHK(s0, hipEventSynchronize(b));
  float f = 0.f;
  SCHK(s0, hipEventElapsedTime(&f, a, b));
  *ms = f;
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return 0;
}

// ------------------------------------------------------------- autotuner
struct Tuning {
  int order[PANGU_KF_COUNT];
  int split;
  int share;
  int pair;  // bit d-1: the sweep along d runs two lanes per face
  int march[2];  // faces per thread of the X2 and X3 sweeps (0: one)
  int tile[PANGU_KF_COUNT];  // cells per tile of each family
  int persist;               // recovery as a persistent work-queue kernel
};
static std::map<std::string, Tuning> g_tune_cache;

// Identifies a launch workload: block shapes and kernel-selecting options.
static std::string tune_signature(const pangu_sim *s) {
  std::string k;
  char buf[96];
  const pangu_sim_cfg &c = s->cfg;
  snprintf(buf, sizeof(buf), "m%d r%d l%d c%d B%d H%d g%d|", c.metric, c.riemann, c.limiter,
           c.recovery, c.enable_B, c.enable_heating, c.ng);
  k += buf;
  int run = 0;
  for (int b = 0; b <= s->nblk; ++b) {
    const bool same = b < s->nblk && b > 0 && s->hb[b].n[0] == s->hb[b - 1].n[0] &&
                      s->hb[b].n[1] == s->hb[b - 1].n[1] && s->hb[b].n[2] == s->hb[b - 1].n[2];
    if (b > 0 && !same) {
      const HostBlock &H = s->hb[b - 1];
      snprintf(buf, sizeof(buf), "%dx%dx%d*%d,", H.n[0], H.n[1], H.n[2], run);
      k += buf;
      run = 0;
    }
    ++run;
  }
  return k;
}

static double family_ms(int fam, const std::vector<std::string> &names, const std::vector<double> &tot) {
  static const char *const fam_kernels[PANGU_KF_COUNT][3] = {
      {"fix_cons", nullptr, nullptr},     {"flux_x1", nullptr, nullptr},
      {"flux_x2", nullptr, nullptr},      {"flux_x3", nullptr, nullptr},
      {"ct_emf", "ct_faces", nullptr},    {"div_update_source", "div_update", "source"},
      {"recovery_heating", "recovery", "heating"}};
  double ms = 0.0;
  for (size_t j = 0; j < names.size(); ++j)
    for (int q = 0; q < 3; ++q)
      if (fam_kernels[fam][q] && names[j] == fam_kernels[fam][q]) ms += tot[j];
  return ms;
}

static int apply_tuning(pangu_sim *s, const Tuning &T) {
  for (int f = 0; f < PANGU_KF_COUNT; ++f) PCHK(s, pangu_set_tile_order(s->ctx, f, T.order[f]));
  PCHK(s, pangu_set_split_source(s->ctx, T.split));
  PCHK(s, pangu_set_share_reconstruction(s->ctx, T.share));
  PCHK(s, pangu_set_pair_flux(s->ctx, T.pair));
  for (int d = 0; d < 2; ++d) PCHK(s, pangu_set_flux_march(s->ctx, d + 2, T.march[d]));
  for (int f = 0; f < PANGU_KF_COUNT; ++f) PCHK(s, pangu_set_tile_size(s->ctx, f, T.tile[f]));
  PCHK(s, pangu_set_persistent_recovery(s->ctx, T.persist));
  drop_graph(s);
  return 0;
}

int pangu_sim_autotune(pangu_sim *s, int cycles, int use_cache, int *out_orders, int *out_split,
                       int *out_share, int *out_pair, int *out_march, int *out_tiles,
                       int *out_persist) {
  if (!s) return PANGU_EINVAL;
  (void)hipSetDevice(s->device);
  if (cycles < 1) cycles = 1;
  const std::string sig = tune_signature(s);
  Tuning best;
  auto hit = g_tune_cache.find(sig);
  if (use_cache && hit != g_tune_cache.end()) {
    best = hit->second;
  } else {
    if (int r = ensure_stage1(s)) return r;
    if (!s->dt_on_device) {
      push_dt_to_device(s);
      s->dt_on_device = true;
    }
    // Snapshot of everything a fused cycle reads or leaves behind.
    struct Save { void *p; size_t bytes; void *copy; };
    const int64_t T = s->total;
    std::vector<Save> save = {
        {s->P[0], (size_t)(T * s->np) * 8, nullptr}, {s->P[1], (size_t)(T * s->np) * 8, nullptr},
        {s->U[0], (size_t)(T * s->nc) * 8, nullptr}, {s->U[1], (size_t)(T * s->nc) * 8, nullptr},
        {s->flag, (size_t)T * 8, nullptr},           {s->flag_rec, (size_t)T * 8, nullptr},
        {s->alfven, (size_t)(T * 3) * 8, nullptr},   {s->emf, (size_t)(T * 3) * 8, nullptr},
        {s->mask, (size_t)T, nullptr},               {s->status, (size_t)T * 4, nullptr},
        {s->iters, (size_t)T * 4, nullptr},          {s->dstate, (size_t)s->nmesh * 16, nullptr}};
    auto release = [&]() {
      for (auto &x : save)
        if (x.copy) (void)hipFree(x.copy);
    };
    for (auto &x : save) {
      if (hipMalloc(&x.copy, x.bytes) != hipSuccess) {
        x.copy = nullptr;
        release();
        return sfail(s, PANGU_ENOMEM, "autotune: no memory for the state snapshot");
      }
      if (hipMemcpyAsync(x.copy, x.p, x.bytes, hipMemcpyDeviceToDevice, s->stream) != hipSuccess) {
        release();
        return sfail(s, PANGU_EDEVICE, "autotune: snapshot copy failed");
      }
    }
    const int ncycle0 = s->ncycle;
    for (int f = 0; f < PANGU_KF_COUNT; ++f) best.order[f] = pangu_get_tile_order(s->ctx, f);
    best.split = 0;
    best.share = 0;
    best.pair = 0;
    best.march[0] = best.march[1] = 0;
    for (int f = 0; f < PANGU_KF_COUNT; ++f) best.tile[f] = 256;
    best.persist = 0;
    int rc = apply_tuning(s, best);
    // coordinate descent: one family at a time, candidates timed with the
    // kernel clock over `cycles` fused cycles
    auto measure = [&](const Tuning &cand, int fam, double *ms) -> int {
      if (int r = apply_tuning(s, cand)) return r;
      std::vector<std::string> names;
      std::vector<double> tot;
      std::vector<int> calls;
      if (int r = clock_cycles(s, cycles, names, tot, calls)) return r;
      *ms = family_ms(fam, names, tot);
      return 0;
    };
    if (!rc) {
      double base_ms = 0, alt_ms = 0;
      Tuning alt = best;
      alt.share = 1;
      rc = measure(best, PANGU_KF_FLUX_X1, &base_ms);
      if (!rc) rc = measure(alt, PANGU_KF_FLUX_X1, &alt_ms);
      if (!rc && alt_ms > 0 && alt_ms < base_ms) best = alt;
    }
    for (int f = 0; f < PANGU_KF_COUNT && !rc; ++f) {
      double base_ms = 0, alt_ms = 0;
      Tuning alt = best;
      alt.order[f] = 1 - best.order[f];
      rc = measure(best, f, &base_ms);
      if (!rc) rc = measure(alt, f, &alt_ms);
      if (!rc && alt_ms > 0 && alt_ms < base_ms) best = alt;
    }
    // workgroup shape of each flux sweep: 256 faces of one lane each, or 128
    // faces of two lanes each (the X1 sweep only when it does not share
    // reconstructions)
    for (int d = 0; d < 3 && !rc; ++d) {
      const int f = PANGU_KF_FLUX_X1 + d;
      if (d == 0 && best.share) continue;
      double base_ms = 0, alt_ms = 0;
      Tuning alt = best;
      alt.pair = best.pair ^ (1 << d);
      rc = measure(best, f, &base_ms);
      if (!rc) rc = measure(alt, f, &alt_ms);
      if (!rc && base_ms > 0 && alt_ms > 0 && alt_ms < base_ms) best = alt;
    }
    // faces per thread of the X2 and X3 sweeps (marching columns; chunks
    // of the sweep's tile order)
    for (int d = 0; d < 2 && !rc; ++d) {
      const int f = PANGU_KF_FLUX_X2 + d;
      static const int depths[] = {2, 3, 4, 6, 8, 12};
      double best_ms = 0;
      rc = measure(best, f, &best_ms);
      Tuning win = best;
      for (int depth : depths) {
        if (rc) break;
        Tuning alt = best;
        alt.march[d] = depth;
        double ms = 0;
        rc = measure(alt, f, &ms);
        if (!rc && ms > 0 && best_ms > 0 && ms < best_ms) {
          best_ms = ms;
          win = alt;
        }
      }
      best = win;
    }
    // cells per tile of each family (the flux sweeps in their one-lane form)
    for (int f = 0; f < PANGU_KF_COUNT && !rc; ++f) {
      if (f >= PANGU_KF_FLUX_X1 && f <= PANGU_KF_FLUX_X3) {
        const int d = f - PANGU_KF_FLUX_X1;
        if ((d == 0 && best.share) || ((best.pair >> d) & 1) || (d > 0 && best.march[d - 1])) continue;
      }
      double best_ms = 0;
      rc = measure(best, f, &best_ms);
      Tuning win = best;
      for (int cells : {128, 64}) {
        if (rc) break;
        Tuning alt = best;
        alt.tile[f] = cells;
        double ms = 0;
        rc = measure(alt, f, &ms);
        if (!rc && ms > 0 && best_ms > 0 && ms < best_ms) {
          best_ms = ms;
          win = alt;
        }
      }
      best = win;
    }
    // recovery: one workgroup per tile, or persistent workgroups pulling
    // tiles from a queue
    if (!rc) {
      double base_ms = 0, alt_ms = 0;
      Tuning alt = best;
      alt.persist = 1;
      rc = measure(best, PANGU_KF_RECOVERY, &base_ms);
      if (!rc) rc = measure(alt, PANGU_KF_RECOVERY, &alt_ms);
      if (!rc && base_ms > 0 && alt_ms > 0 && alt_ms < base_ms) best = alt;
    }
    if (!rc) {
      double base_ms = 0, alt_ms = 0;
      Tuning alt = best;
      alt.split = 1;
      rc = measure(best, PANGU_KF_DIV_UPDATE, &base_ms);
      if (!rc) rc = measure(alt, PANGU_KF_DIV_UPDATE, &alt_ms);
      if (!rc && alt_ms > 0 && alt_ms < base_ms) best = alt;
    }
    for (auto &x : save)
      if (hipMemcpyAsync(x.p, x.copy, x.bytes, hipMemcpyDeviceToDevice, s->stream) != hipSuccess)
        rc = rc ? rc : sfail(s, PANGU_EDEVICE, "autotune: restore copy failed");
    (void)hipStreamSynchronize(s->stream);
    release();
    s->ncycle = ncycle0;
    if (rc) return rc;
    g_tune_cache[sig] = best;
  }
  if (int r = apply_tuning(s, best)) return r;
  if (out_orders)
    for (int f = 0; f < PANGU_KF_COUNT; ++f) out_orders[f] = best.order[f];
  if (out_split) *out_split = best.split;
  if (out_share) *out_share = best.share;
  if (out_pair) *out_pair = best.pair;
  if (out_march)
    for (int d = 0; d < 2; ++d) out_march[d] = best.march[d];
  if (out_tiles)
    for (int f = 0; f < PANGU_KF_COUNT; ++f) out_tiles[f] = best.tile[f];
  if (out_persist) *out_persist = best.persist;
  return 0;
}

int pangu_sim_set_tile_order(pangu_sim *s, int fam, int order) {
  PCHK(s, pangu_set_tile_order(s->ctx, fam, order));
  drop_graph(s);
  return 0;
}

int pangu_sim_set_c2p_guess(pangu_sim *s, pangu_mlp *m, double delta) {
  PCHK(s, pangu_set_c2p_guess(s->ctx, m, delta));
  drop_graph(s);
  return 0;
}

int pangu_sim_set_c2p_gate(pangu_sim *s, int min_prev_iters) {
  PCHK(s, pangu_set_c2p_gate(s->ctx, min_prev_iters));
  drop_graph(s);
  return 0;
}

int pangu_sim_set_share_reconstruction(pangu_sim *s, int share) {
  PCHK(s, pangu_set_share_reconstruction(s->ctx, share));
  drop_graph(s);
  return 0;
}

int pangu_sim_set_pair_flux(pangu_sim *s, int dir_mask) {
  PCHK(s, pangu_set_pair_flux(s->ctx, dir_mask));
  drop_graph(s);
  return 0;
}

int pangu_sim_set_persistent_recovery(pangu_sim *s, int on) {
  PCHK(s, pangu_set_persistent_recovery(s->ctx, on));
  drop_graph(s);
  return 0;
}

int pangu_sim_set_tile_size(pangu_sim *s, int kernel_family, int cells) {
  PCHK(s, pangu_set_tile_size(s->ctx, kernel_family, cells));
  drop_graph(s);
  return 0;
}

int pangu_sim_set_flux_march(pangu_sim *s, int dir, int depth) {
  PCHK(s, pangu_set_flux_march(s->ctx, dir, depth));
  drop_graph(s);
  return 0;
}

int pangu_sim_set_split_source(pangu_sim *s, int split) {
  PCHK(s, pangu_set_split_source(s->ctx, split));
  drop_graph(s);
  return 0;
}

int pangu_sim_set_fuse_ct(pangu_sim *s, int fuse) {
  PCHK(s, pangu_set_fuse_ct(s->ctx, fuse));
  drop_graph(s);
  return 0;
}

int pangu_sim_set_fuse_fix_x1(pangu_sim *s, int on) {
  PCHK(s, pangu_set_fuse_fix_x1(s->ctx, on));
  drop_graph(s);
  return 0;
}

int pangu_sim_set_fuse_update_recovery(pangu_sim *s, int on) {
  PCHK(s, pangu_set_fuse_update_recovery(s->ctx, on));
  drop_graph(s);
  return 0;
}

int pangu_sim_set_overlap_exchange(pangu_sim *s, int on) {
  PCHK(s, pangu_set_overlap_exchange(s->ctx, on));
  drop_graph(s);
  return 0;
}

const char *pangu_sim_kernel_name(pangu_sim *s, int n) {
  return (s && n >= 0 && n < (int)s->knames.size()) ? s->knames[n].c_str() : "";
}

static const char *kPhase[] = {"fix_cons",  "fluxes",  "ct",  "div_update_source",
                               "exchange", "recovery_heating", "fix_recovery", "bc", "timestep"};

const char *pangu_sim_phase_name(int n) {
  return (n >= 0 && n < (int)(sizeof(kPhase) / sizeof(kPhase[0]))) ? kPhase[n] : "";
}

int pangu_sim_time_phases(pangu_sim *s, int reps, double *ms_out, int maxp) {
  SCHK(s, hipSetDevice(s->device));
  if (int r = ensure_stage1(s)) return r;
  if (!s->dt_on_device) {
    push_dt_to_device(s);
    s->dt_on_device = true;
  }
  const int nph = (int)(sizeof(kPhase) / sizeof(kPhase[0]));
  pangu_ctx *x = s->ctx;
  void *st = s->stream;
  const int c0 = s->h_s1, c1 = s->h_base, ci = s->h_base;
  for (int ph = 0; ph < nph && ph < maxp; ++ph) {
    hipEvent_t a, b;
    SCHK(s, hipEventCreate(&a));
    SCHK(s, hipEventCreate(&b));
    SCHK(s, hipEventRecord(a, s->stream));
    for (int r = 0; r < reps; ++r) {
      switch (ph) {
        case 0: PCHK(s, pangu_fix_primitive(x, c0, st)); PCHK(s, pangu_calculate_conservative(x, c0, st)); break;
        case 1: PCHK(s, pangu_calculate_fluxes(x, c0, st)); break;
        case 2: if (s->cfg.enable_B) PCHK(s, pangu_constrained_transport(x, c0, st)); break;
        case 3: PCHK(s, pangu_flux_divergence(x, c0, s->h_dudt, st));
                PCHK(s, pangu_update_independent(x, ci, s->h_dudt, 0.0, c1, st));
                PCHK(s, pangu_add_geometric_source(x, c1, 0.0, st)); break;
        case 4: PCHK(s, pangu_exchange_ghosts(x, c1, st)); break;
        case 5: PCHK(s, pangu_recovery(x, c1, st)); if (s->cfg.enable_heating) PCHK(s, pangu_apply_electron_heating(x, c1, st)); break;
        case 6: if (s->cfg.recovery != PANGU_RECOVERY_ROBUSTER) PCHK(s, pangu_fix_recovery(x, c1, st)); break;
        case 7: PCHK(s, pangu_apply_boundary_conditions(x, c1, st)); break;
        case 8: { std::vector<double> e(s->nmesh); PCHK(s, pangu_estimate_timestep(x, c1, e.data(), st)); } break;
      }
    }
    SCHK(s, hipEventRecord(b, s->stream));
    SCHK(s, hipEventSynchronize(b));
    float f = 0.f;
    SCHK(s, hipEventElapsedTime(&f, a, b));
    ms_out[ph] = (double)f / reps;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
  }
  return nph;
}

int pangu_sim_task(pangu_sim *s, int task, int c0i, int c1i, double bdt, double *out) {
  SCHK(s, hipSetDevice(s->device));
  if (int r = ensure_stage1(s)) return r;
  const int c0 = c0i == 0 ? s->h_base : s->h_s1;
  const int c1 = c1i == 0 ? s->h_base : s->h_s1;
  pangu_ctx *x = s->ctx;
  void *st = s->stream;
  switch (task) {
    case T_FIX: PCHK(s, pangu_fix_primitive(x, c0, st)); break;
    case T_CONS: PCHK(s, pangu_calculate_conservative(x, c0, st)); break;
    case T_FLUXES: PCHK(s, pangu_calculate_fluxes(x, c0, st)); break;
    case T_CT: PCHK(s, pangu_constrained_transport(x, c0, st)); break;
    case T_DIV: PCHK(s, pangu_flux_divergence(x, c0, s->h_dudt, st)); break;
    case T_UPDATE: PCHK(s, pangu_update_independent(x, s->h_base, s->h_dudt, bdt, c1, st)); break;
    case T_SOURCE: PCHK(s, pangu_add_geometric_source(x, c1, bdt, st)); break;
    case T_RECOVERY: PCHK(s, pangu_recovery(x, c1, st)); break;
    case T_HEATING: PCHK(s, pangu_apply_electron_heating(x, c1, st)); break;
    case T_FIXREC: PCHK(s, pangu_fix_recovery(x, c1, st)); break;
    case T_EXCHANGE: PCHK(s, pangu_exchange_ghosts(x, c1, st)); break;
    case T_BC: PCHK(s, pangu_apply_boundary_conditions(x, c1, st)); break;
    case T_FLUXCORR: PCHK(s, pangu_flux_correction(x, c0, st)); break;
    case T_TIMESTEP: {
      std::vector<double> e(s->nmesh);
      PCHK(s, pangu_estimate_timestep(x, c1, e.data(), st));
      if (out) *out = e[0];
    } break;
    default: return sfail(s, PANGU_EINVAL, "unknown task");
  }
  SCHK(s, hipStreamSynchronize(s->stream));
  return 0;
}

static void *field_ptr(pangu_sim *s, int f, int c, int64_t *bytes) {
  const int64_t T = s->total;
  switch (f) {
    case F_PRIM: *bytes = T * s->np * 8; return s->P[c];
    case F_CONS: *bytes = T * s->nc * 8; return s->U[c];
    case F_FLUX1: *bytes = T * s->nc * 8; return s->F[0];
    case F_FLUX2: *bytes = T * s->nc * 8; return s->F[1];
    case F_FLUX3: *bytes = T * s->nc * 8; return s->F[2];
    case F_ALFVEN: *bytes = T * 3 * 8; return s->alfven;
    case F_EMF: *bytes = T * 3 * 8; return s->emf;
    case F_FLAG: *bytes = T * 8; return s->flag;
    case F_MASK: *bytes = T; return s->mask;
    case F_STATUS: *bytes = T * 4; return s->status;
    case F_ITERS: *bytes = T * 4; return s->iters;
    case F_FLAG_REC: *bytes = T * 8; return s->flag_rec;
    case F_DUDT: *bytes = T * s->nc * 8; return s->dudt;
    case F_ROOT: *bytes = s->root ? T * 8 : 0; return s->root;
  }
  *bytes = 0;
  return nullptr;
}

int64_t pangu_sim_field_size(pangu_sim *s, int f) {
  int64_t b = 0;
  field_ptr(s, f, 0, &b);
  return b;
}

int pangu_sim_get(pangu_sim *s, int f, int c, void *out) {
  int64_t b = 0;
  void *p = field_ptr(s, f, c, &b);
  if (!p) return PANGU_EINVAL;
  SCHK(s, hipStreamSynchronize(s->stream));
  SCHK(s, hipMemcpy(out, p, (size_t)b, hipMemcpyDeviceToHost));
  return 0;
}

int pangu_sim_set(pangu_sim *s, int f, int c, const void *in) {
  int64_t b = 0;
  void *p = field_ptr(s, f, c, &b);
  if (!p) return PANGU_EINVAL;
  SCHK(s, hipStreamSynchronize(s->stream));
  SCHK(s, hipMemcpy(p, in, (size_t)b, hipMemcpyHostToDevice));
  s->stage1_ready = true;
  return 0;
}

double pangu_sim_time(pangu_sim *s) {
  (void)hipStreamSynchronize(s->stream);
  sync_dt_from_device(s);
  return s->time;
}
double pangu_sim_dt(pangu_sim *s) {
  (void)hipStreamSynchronize(s->stream);
  sync_dt_from_device(s);
  return s->dt;
}

int pangu_sim_mesh_state(pangu_sim *s, int m, double *dt, double *time, int *first_block,
                         int *nblocks) {
  if (!s || m < 0 || m >= s->nmesh) return PANGU_EINVAL;
  (void)hipStreamSynchronize(s->stream);
  sync_dt_from_device(s);
  int first = 0;
  while (s->bmesh[first] != m) ++first;
  int last = first;
  while (last < s->nblk && s->bmesh[last] == m) ++last;
  if (dt) *dt = s->mdt[m];
  if (time) *time = s->mtime[m];
  if (first_block) *first_block = first;
  if (nblocks) *nblocks = last - first;
  return 0;
}

int pangu_sim_nblocks(pangu_sim *s) { return s ? s->nblk : 0; }

int pangu_sim_diagnostics(pangu_sim *s, int container, const int *shell, double *out) {
  if (!s || !shell || !out || container < 0 || container > 1) return PANGU_EINVAL;
  if (!s->ddiag) SCHK(s, hipMalloc((void **)&s->ddiag, (size_t)s->nmesh * PANGU_NDIAG * sizeof(double)));
  PCHK(s, pangu_diagnostics(s->ctx, container == 0 ? s->h_base : s->h_s1, shell, s->ddiag, s->stream));
  SCHK(s, hipMemcpy(out, s->ddiag, (size_t)s->nmesh * PANGU_NDIAG * sizeof(double), hipMemcpyDeviceToHost));
  return 0;
}

int pangu_sim_record_roots(pangu_sim *s, int on) {
  if (!s) return PANGU_EINVAL;
  SCHK(s, hipStreamSynchronize(s->stream));
  if (on && !s->root) {
    if (dalloc(s, &s->root, s->total)) return PANGU_ENOMEM;
  } else if (!on && s->root) {
    (void)hipFree(s->root);
    s->root = nullptr;
  }
  if (int r = register_shared(s)) return sfail(s, r, pangu_last_error(s->ctx));
  drop_graph(s);  // a captured cycle holds the old tables
  return 0;
}

int pangu_sim_context(pangu_sim *s, pangu_ctx **ctx, int *handles, void **stream) {
  if (!s || !ctx || !handles || !stream) return PANGU_EINVAL;
  *ctx = s->ctx;
  handles[0] = s->h_base;
  handles[1] = s->h_s1;
  *stream = (void *)s->stream;
  return 0;
}

int pangu_sim_block_geom(pangu_sim *s, int b, double *xd, int *n) {
  if (b < 0 || b >= s->nblk) return PANGU_EINVAL;
  const HostBlock &H = s->hb[b];
  for (int d = 0; d < 3; ++d) {
    xd[d] = H.xg[d];
    xd[3 + d] = H.dx[d];
    n[d] = H.n[d];
    n[3 + d] = H.s[d];
    n[6 + d] = H.e[d];
  }
  return 0;
}

int pangu_sim_block_loc(pangu_sim *s, int b, int *out) {
  if (!s || b < 0 || b >= s->nblk) return PANGU_EINVAL;
  out[0] = s->hb[b].level;
  for (int d = 0; d < 3; ++d) out[1 + d] = s->hb[b].lx[d];
  return 0;
}

const char *pangu_sim_last_error(pangu_sim *s) { return s ? s->err.c_str() : "null"; }

// Plain device copy of n doubles, used to calibrate the HBM byte counters
// (FETCH_SIZE / WRITE_SIZE) against a kernel of known traffic.
__global__ void k_copy_calibration(const double *__restrict__ a, double *__restrict__ b, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

int pangu_sim_copy_calibration(pangu_sim *s, int64_t n, int reps, double *ms) {
  double *a = nullptr, *b = nullptr;
  SCHK(s, hipMalloc((void **)&a, (size_t)n * sizeof(double)));
  SCHK(s, hipMalloc((void **)&b, (size_t)n * sizeof(double)));
  SCHK(s, hipMemsetAsync(a, 0, (size_t)n * sizeof(double), s->stream));
  hipEvent_t e0, e1;
  SCHK(s, hipEventCreate(&e0));
  SCHK(s, hipEventCreate(&e1));
  const int64_t nblk = (n + 255) / 256;
  hipLaunchKernelGGL(k_copy_calibration, dim3((unsigned)nblk), dim3(256), 0, s->stream, a, b, n);
  SCHK(s, hipEventRecord(e0, s->stream));
  for (int r = 0; r < reps; ++r)
    hipLaunchKernelGGL(k_copy_calibration, dim3((unsigned)nblk), dim3(256), 0, s->stream, a, b, n);
  SCHK(s, hipEventRecord(e1, s->stream));
  SCHK(s, hipEventSynchronize(e1));
  float f = 0.f;
  SCHK(s, hipEventElapsedTime(&f, e0, e1));
  *ms = (double)f / reps;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  return 0;
}

}  // extern "C"
