// Oracle mesh/driver state.  TEST INFRASTRUCTURE ONLY.
#pragma once
#include <array>
#include <map>
#include <vector>

#include "oracle.h"
#include "oracle_api.h"

namespace oracle {

struct Block {
  int nx[3], ng[3], n[3], s[3], e[3], lx[3], gofs[3], level;
  double xmin[3], xmax[3], dx[3], xg[3];
  int64_t cells, off;
  int64_t idx(int k, int j, int i) const { return ((int64_t)k * n[1] + j) * n[0] + i; }
};

// Parthenon UniformCartesian: Xc = xg + (idx + 0.5) dx, Xf = xg + idx dx
// with xg the ghost-inclusive origin.
static inline double Xc(const Block &B, int d, int idx) { return B.xg[d] + (idx + 0.5) * B.dx[d]; }
static inline double Xf(const Block &B, int d, int idx) { return B.xg[d] + idx * B.dx[d]; }

// Edge idx of n equal cells over [xmin, xmax] (Parthenon's uniform mesh
// generator; the mesh edges are exact).
double mesh_edge(int64_t idx, int64_t n, double xmin, double xmax);

struct Geo {
  real gcov[4][4], gcon[4][4], gdet;
  int ok;
};
struct BlockMemo {
  int ni, nj;
  std::vector<Geo> cen, f1, f2;
  std::vector<double> conn;
};

struct Sim {
  oracle_cfg cfg;
  Params p;
  int ndim = 0, np = 0, nc = 0, iENT = 0, iKEL = -1, iBX = -1;
  int slot_of[NPRIM];
  int nblk = 0;
  std::vector<Block> blk;
  int64_t total_cells = 0;
  std::vector<double> P[2], U[2], dudt, F[3], alfven, emf, flag, flag_rec;
  std::vector<uint8_t> mask;
  std::vector<int32_t> c2p_status, c2p_iters;
  bool stage1_ready = false;
  double time = 0, dt = 0, dt_est = 0;
  int ncycle = 0;
  bool memo = false;
  std::vector<BlockMemo> gm;

  double *Pc(int c, int b, int q) { return P[c].data() + blk[b].off * np + (int64_t)q * blk[b].cells; }
  const double *Pc(int c, int b, int q) const {
    return P[c].data() + blk[b].off * np + (int64_t)q * blk[b].cells;
  }
  double *Uc(int c, int b, int q) { return U[c].data() + blk[b].off * nc + (int64_t)q * blk[b].cells; }
  double *Fc(int d, int b, int q) { return F[d].data() + blk[b].off * nc + (int64_t)q * blk[b].cells; }
  double *Ec(int b, int d) { return emf.data() + blk[b].off * 3 + (int64_t)d * blk[b].cells; }

  void setup(const oracle_cfg &c);
  // static mesh refinement (refine.cpp)
  bool refined = false;
  std::map<std::array<int64_t, 4>, int> leaf_of;  // (level, lx1, lx2, lx3) -> block
  void make_leaves(std::vector<std::array<int64_t, 4>> &out) const;
  int leaf_at(int level, const int64_t g[3]) const;
  bool wrap_cell(int d, int level, int64_t &g) const;
  double level_value(int c, int q, int level, const int64_t g[3]) const;
  void exchange_refined(int c);
  void build_geometry_memo();
  int metric(int b, int loc, int k, int j, int i, real gcov[4][4], real gcon[4][4], real *gdet) const;
  void connection(int b, int k, int j, int i, real conn[4][4][4]) const;
  void load_prim(int c, int b, int64_t id, real pc[NPRIM]) const;

  void fix_primitive(int c);
  void calc_conservative(int c);
  void sweep(int c, int dir);
  void fluxes(int c);
  void ct();
  void flux_div();
  void update(double bdt, int out);
  void source(int c, double bdt);
  void recovery(int c);
  void heating(int c);
  void fix_recovery(int c);
  void exchange(int c);
  void flux_correction();
  void apply_bc(int c);
  double estimate_dt() const;
  void diagnostics(int c, int shell, double out[ODIAG_N], double mag[ODIAG_N]) const;

  void ensure_stage_container();
  void run_stage(int stage);
  void cycle();
  int init_problem();
};

}  // namespace oracle
