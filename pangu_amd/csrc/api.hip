// C ABI of pangu_hip: context, geometry tables, container registration and
// the task entry points (see include/pangu_hip.h).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/pangu_hip.h"
#include "../../include/pangu_mlp.h"
#include "common.h"
#include "host.h"
#include "kernels.h"

using namespace pangu;

struct DevContainer {
  std::vector<double *> hprim, hcons, hflux[3];
  double **dprim = nullptr, **dcons = nullptr, **dflux[3] = {nullptr, nullptr, nullptr};
  KFields kf() const {
    KFields f;
    f.prim = dprim;
    f.cons = dcons;
    for (int d = 0; d < 3; ++d) f.flux[d] = dflux[d];
    return f;
  }
};

struct pangu_ctx {
  pangu_config cfg;                 // physics shared by every mesh (mesh 0's)
  std::vector<pangu_config> mcfg;   // per mesh: metric parameters, topology, boundaries
  std::vector<int> bmesh;           // mesh of each block
  std::vector<int> mfirst;          // first block of each mesh (nmesh + 1)
  int device = 0;
  int nmesh = 1, nplanes = 0, ndt = 0;
  int nblocks = 0, ndim = 1, np = 0, nc = 0, iBX = -1, iENT = 0, iKEL = -1;
  std::vector<HostBlock> hb;
  std::vector<int *> tab_bufs;      // launch tables
  DevBlock *dblk = nullptr;
  DevGeo *dgeo = nullptr;
  std::vector<double *> geo_bufs;
  int *dnbr = nullptr;
  // static mesh refinement: gather tables (L.ml points into these)
  XFill *dxfill = nullptr;
  XProl *dxprol = nullptr;
  XFlux *dxflux = nullptr;
  std::vector<DevContainer> conts;
  // shared tables
  double **dalfven = nullptr, **demf = nullptr, **dflag = nullptr, **dflag_rec = nullptr;
  uint8_t **dmask = nullptr;
  int32_t **dstatus = nullptr, **diters = nullptr;
  double **droot = nullptr;
  bool shared_ok = false;
  // learned Kastaun bracket (pangu_set_c2p_guess, pangu_set_c2p_gate)
  pangu_mlp *mlp = nullptr;
  double guess_delta = 0.0;
  int guess_gate = 0;
  // scratch
  double *dred = nullptr;
  int *dcluster = nullptr;
  unsigned *dqueue = nullptr;  // work queue of the persistent recovery kernel
  double *dest = nullptr;
  // diagnostics: per-tile partial sums, per-mesh tile ranges, shell indices
  double *ddiag_part = nullptr;
  int *ddiag_mtile = nullptr, *ddiag_shell = nullptr;
  std::vector<int> diag_mtile;         // first interior tile of each mesh (nmesh + 1)
  // the device's static diagnostics workspace (pangu_diagnostics_noalloc)
  double *dscratch = nullptr;
  std::string scratch_err;
  // second stream of the fused stage (default priority) and its fork / join
  // events: the ghost exchange of the output container runs there
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  LaunchCtx L;
  std::string err;
};

static int fail(pangu_ctx *c, int code, const std::string &msg) {
  if (c) c->err = msg;
  return code;
}

#define HIPCHK(ctx, call)                                                              \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess)                                                              \
      return fail(ctx, PANGU_EDEVICE, std::string(#call ": ") + hipGetErrorString(e_)); \
  } while (0)

template <class T>
static int upload(pangu_ctx *c, T **dst, const T *src, size_t n) {
  HIPCHK(c, hipMalloc((void **)dst, n * sizeof(T) > 0 ? n * sizeof(T) : 1));
  if (n) HIPCHK(c, hipMemcpy(*dst, src, n * sizeof(T), hipMemcpyHostToDevice));
  return 0;
}

static int validate(const pangu_config *cfg, const pangu_block *blocks, int nb, std::string &msg) {
  if (!cfg || !blocks || nb <= 0) {
    msg = "null configuration or no blocks";
    return 0;
  }
  if (cfg->metric < 0 || cfg->metric > 3) { msg = "metric must be minkowski, bl, cks or mks"; return 0; }
  if (cfg->riemann != PANGU_RIEMANN_LAXF && cfg->riemann != PANGU_RIEMANN_HLL) {
    msg = "Unknown Riemann solver (laxf or hll)";
    return 0;
  }
  if (cfg->limiter != PANGU_LIMITER_MC && cfg->limiter != PANGU_LIMITER_PPM4) {
    msg = "limiter must be mc or ppm4";
    return 0;
  }
  if (cfg->recovery < 0 || cfg->recovery > 2) {
    msg = "<core>/recovery must be 'chainer' or 'robuster' (or onedvsq)";
    return 0;
  }
  if (cfg->heat_model < 0 || cfg->heat_model > 5) {
    msg = "Invalid electrons/model. Allowed values: constant, howes, kawazura, werner, rowan, sharma.";
    return 0;
  }
  if (!(cfg->gamma > 1.0)) { msg = "adiabatic_index must exceed 1"; return 0; }
  for (int b = 0; b < nb; ++b)
    for (int d = 0; d < 3; ++d) {
      if (blocks[b].nx[d] != blocks[0].nx[d] || blocks[b].ng[d] != blocks[0].ng[d]) {
        msg = "all meshblocks of a mesh must share one shape";
        return 0;
      }
      if (blocks[b].level < 0 || blocks[b].level > 20) {
        msg = "refinement level out of range (0..20)";
        return 0;
      }
      if (blocks[b].nx[d] < 1 ||
          (blocks[b].nx[d] > 1 && (int64_t)blocks[b].lx[d] >= ((int64_t)cfg->nb[d] << blocks[b].level)) ||
          blocks[b].lx[d] < 0) {
        msg = "block extent or logical location out of range";
        return 0;
      }
      if (blocks[b].nx[d] > 1 && blocks[b].ng[d] < (cfg->limiter == PANGU_LIMITER_PPM4 ? 3 : 2)) {
        msg = "nghost too small for the limiter (ppm4 reads i-3..i+2)";
        return 0;
      }
    }
  return 1;
}

// Every mesh of a batched context runs the same kernels: only the metric
// parameters (a, h), the block topology (nb) and the boundaries may differ.
static int same_physics(const pangu_config &x, const pangu_config &y) {
  return x.metric == y.metric && x.gamma == y.gamma && x.cfl == y.cfl && x.riemann == y.riemann &&
         x.limiter == y.limiter && x.recovery == y.recovery && x.enable_B == y.enable_B &&
         x.enable_heating == y.enable_heating && x.density_floor == y.density_floor &&
         x.energy_floor == y.energy_floor && x.density_floor_pow == y.density_floor_pow &&
         x.energy_floor_pow == y.energy_floor_pow && x.sigma_max == y.sigma_max &&
         x.lorentz_max == y.lorentz_max && x.heat_model == y.heat_model &&
         x.fel_constant == y.fel_constant && x.gamma_e == y.gamma_e && x.gamma_p == y.gamma_p &&
         x.ratio_min == y.ratio_min && x.ratio_max == y.ratio_max && x.limit_kel == y.limit_kel &&
         x.suppress_highb_heat == y.suppress_highb_heat &&
         x.enforce_positive_dissipation == y.enforce_positive_dissipation && x.excise == y.excise &&
         x.r_excise == y.r_excise && x.dexcise == y.dexcise && x.pexcise == y.pexcise;
}

static int validate_batch(const pangu_config *cfgs, int nmesh, const pangu_block *blocks,
                          const int *bmesh, int nblocks, std::string &msg) {
  if (!cfgs || nmesh <= 0 || !blocks || nblocks < nmesh) {
    msg = "need at least one block per mesh";
    return 0;
  }
  int first = 0;
  for (int m = 0; m < nmesh; ++m) {
    int last = first;
    while (last < nblocks && (bmesh ? bmesh[last] : 0) == m) ++last;
    if (last == first) {
      msg = "blocks must be grouped by mesh in increasing mesh order, each mesh non-empty";
      return 0;
    }
    if (!validate(&cfgs[m], blocks + first, last - first, msg)) return 0;
    if (m > 0 && !same_physics(cfgs[0], cfgs[m])) {
      msg = "meshes of one context must share every physics parameter except a, h, nb and bc";
      return 0;
    }
    first = last;
  }
  if (first != nblocks) {
    msg = "block mesh indices out of range";
    return 0;
  }
  return 1;
}

// Geometry sets: blocks with identical x1/x2 footprints (and, for x3-dependent
// metrics, identical x3 footprints) and identical metric parameters share one
// table set.
static int build_geometry(pangu_ctx *c, const pangu_block *blocks) {
  std::map<std::vector<double>, int> key2geo;
  std::vector<DevGeo> geos;
  std::vector<DevBlock> db(c->nblocks);
  for (int b = 0; b < c->nblocks; ++b) {
    const pangu_config &cfg = c->mcfg[c->bmesh[b]];
    const HostBlock &H = c->hb[b];
    const bool has3d = (cfg.metric == PANGU_METRIC_CKS) && H.n[2] > 1;
    DevBlock &D = db[b];
    std::memset(&D, 0, sizeof(D));
    for (int d = 0; d < 3; ++d) {
      D.n[d] = H.n[d];
      D.s[d] = H.s[d];
      D.e[d] = H.e[d];
      D.lx[d] = H.lx[d];
      D.dx[d] = H.dx[d];
      const bool lo_edge = H.lx[d] == 0, hi_edge = (int64_t)H.lx[d] == ((int64_t)cfg.nb[d] << H.level) - 1;
      D.bcf[d][0] = (H.n[d] > 1 && lo_edge && cfg.bc[d][0] != PANGU_BC_PERIODIC) ? cfg.bc[d][0] : -1;
      D.bcf[d][1] = (H.n[d] > 1 && hi_edge && cfg.bc[d][1] != PANGU_BC_PERIODIC) ? cfg.bc[d][1] : -1;
    }
    D.ncell = H.cells;
    D.mesh = c->bmesh[b];
    D.level = H.level;
    std::vector<double> key = {(double)cfg.metric, cfg.a, cfg.h, H.xg[0], H.dx[0], (double)H.n[0],
                               H.xg[1], H.dx[1], (double)H.n[1], (double)H.n[2]};
    if (has3d) {
      key.push_back(H.xg[2]);
      key.push_back(H.dx[2]);
    } else {
      key.push_back(H.xc(2, H.s[2]));
    }
    // floors depend on the mesh's floor parameters, which all meshes share
    auto it = key2geo.find(key);
    if (it != key2geo.end()) {
      D.geo = it->second;
      continue;
    }
    DevGeo G;
    std::memset(&G, 0, sizeof(G));
    G.ni = H.n[0];
    G.nj = H.n[1];
    G.nk = has3d ? H.n[2] : 1;
    G.has3d = has3d ? 1 : 0;
    G.ncen = (int64_t)G.ni * G.nj * G.nk;
    // the kernels address a cell of a field array, a table entry and the
    // 64-slot connection block of a table through 32-bit byte offsets
    if (G.ncen * 64 * (int64_t)sizeof(double) >= ((int64_t)1 << 32) ||
        (int64_t)H.cells * (int64_t)sizeof(double) >= ((int64_t)1 << 32))
      return fail(c, PANGU_EINVAL, "block too large: its geometry table would exceed 2^23 entries");
    std::vector<double> cen((size_t)(G.ncen * CG_NCOMP));
    geom::fill_centre_table(cfg, H, G.ni, G.nj, G.nk, has3d, cen.data());
    // Compact Christoffel storage when the table is symmetric in its lower
    // indices bit for bit (checked for every entry); the source term then
    // loads 40 values per cell instead of 64 and sums the same terms.
    G.conn_sym = 1;
    for (int64_t e = 0; e < G.ncen && G.conn_sym; ++e)
      for (int a = 0; a < 4 && G.conn_sym; ++a)
        for (int b = 0; b < 4; ++b)
          for (int cc = b + 1; cc < 4; ++cc) {
            const double x = cen[(size_t)((CG_CONN + 16 * a + 4 * b + cc) * G.ncen + e)];
            const double y = cen[(size_t)((CG_CONN + 16 * a + 4 * cc + b) * G.ncen + e)];
            if (std::memcmp(&x, &y, sizeof(double)) != 0) G.conn_sym = 0;
          }
    if (G.conn_sym) {
      std::vector<double> compact((size_t)(40 * G.ncen));
      for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b)
          for (int cc = b; cc < 4; ++cc)
            std::memcpy(&compact[(size_t)((10 * a + conn_pair(b, cc)) * G.ncen)],
                        &cen[(size_t)((CG_CONN + 16 * a + 4 * b + cc) * G.ncen)],
                        sizeof(double) * (size_t)G.ncen);
      std::memcpy(&cen[(size_t)(CG_CONN * G.ncen)], compact.data(), compact.size() * sizeof(double));
    }
    double *dcen = nullptr;
    if (upload(c, &dcen, cen.data(), cen.size())) return PANGU_EDEVICE;
    c->geo_bufs.push_back(dcen);
    G.cen = dcen;
    const int nfd = has3d ? 3 : 2;
    const int bdim = H.n[2] > 1 ? 3 : (H.n[1] > 1 ? 2 : 1);
    for (int d = 0; d < 3; ++d) {
      G.fi[d] = G.ni + (d == 0 ? 1 : 0);
      G.fj[d] = G.nj + (d == 1 ? 1 : 0);
      if (d >= bdim || d >= nfd) {
        G.face[d] = nullptr;
        G.nface[d] = 0;
        continue;
      }
      const int fk = G.nk + (d == 2 ? 1 : 0);
      G.nface[d] = (int64_t)G.fi[d] * G.fj[d] * fk;
      std::vector<double> tab((size_t)(G.nface[d] * FG_NCOMP));
      geom::fill_face_table(cfg, H, d, G.fi[d], G.fj[d], fk, has3d, tab.data());
      double *dt = nullptr;
      if (upload(c, &dt, tab.data(), tab.size())) return PANGU_EDEVICE;
      c->geo_bufs.push_back(dt);
      G.face[d] = dt;
    }
    D.geo = (int)geos.size();
    key2geo[key] = D.geo;
    geos.push_back(G);
  }
  if (upload(c, &c->dgeo, geos.data(), geos.size())) return PANGU_EDEVICE;
  if (upload(c, &c->dblk, db.data(), db.size())) return PANGU_EDEVICE;
  return 0;
}

// Neighbour table for same-level exchange: 27 entries per block
// (offset (o0,o1,o2) in {-1,0,1}^3), -1 where a physical boundary is.
// Blocks only neighbour blocks of their own mesh.
static int build_neighbours(pangu_ctx *c) {
  std::vector<int> nbr((size_t)c->nblocks * 27, -1);
  std::map<std::vector<int>, int> at;
  for (int b = 0; b < c->nblocks; ++b)
    at[{c->bmesh[b], c->hb[b].lx[0], c->hb[b].lx[1], c->hb[b].lx[2]}] = b;
  for (int b = 0; b < c->nblocks; ++b) {
    const pangu_config &cfg = c->mcfg[c->bmesh[b]];
    for (int o2 = -1; o2 <= 1; ++o2)
      for (int o1 = -1; o1 <= 1; ++o1)
        for (int o0 = -1; o0 <= 1; ++o0) {
          const int o[3] = {o0, o1, o2};
          std::vector<int> key(4);
          key[0] = c->bmesh[b];
          bool ok = true;
          for (int d = 0; d < 3; ++d) {
            int l = c->hb[b].lx[d] + o[d];
            if (l < 0 || l >= cfg.nb[d]) {
              if (cfg.bc[d][0] == PANGU_BC_PERIODIC)
                l = (l + cfg.nb[d]) % cfg.nb[d];
              else
                ok = false;
            }
            key[d + 1] = l;
          }
          if (!ok) continue;
          auto it = at.find(key);
          if (it != at.end()) nbr[(size_t)b * 27 + (o2 + 1) * 9 + (o1 + 1) * 3 + (o0 + 1)] = it->second;
        }
  }
  return upload(c, &c->dnbr, nbr.data(), nbr.size());
}

// Static mesh refinement, on the host before any device work: checks that
// the blocks tile each mesh and resolves the gather tables of the
// cross-level exchange and of flux correction ("" or why the mesh cannot
// be handled).
struct RefineTables {
  std::vector<XFill> fill;
  std::vector<XProl> prol;
  std::vector<XFlux> flux;
};
static std::string plan_multilevel(const std::vector<HostBlock> &hb, const std::vector<int> &bmesh,
                                   const std::vector<pangu_config> &mcfg, RefineTables &T) {
  if (!refined_mesh(hb)) return "";
  std::string why = check_tiling(hb, bmesh, mcfg);
  if (!why.empty()) return why;
  why = build_refinement_tables(hb, bmesh, mcfg, T.fill, T.prol, T.flux);
  if (!why.empty()) return why;
  if (T.fill.size() > (size_t)0x7fffff00 || T.prol.size() > (size_t)0x7fffff00)
    return "refinement tables exceed 2^31 entries";
  return "";
}

// Uploads the refinement tables (nothing on a single-level mesh).
static int build_multilevel(pangu_ctx *c, const RefineTables &T) {
  if (!refined_mesh(c->hb)) return 0;
  const std::vector<XFill> &fill = T.fill;
  const std::vector<XProl> &prol = T.prol;
  const std::vector<XFlux> &flux = T.flux;
  if (upload(c, &c->dxfill, fill.data(), fill.size()) || upload(c, &c->dxprol, prol.data(), prol.size()) ||
      upload(c, &c->dxflux, flux.data(), flux.size()))
    return PANGU_EDEVICE;
  MultiLevel &M = c->L.ml;
  M.on = 1;
  M.fill = c->dxfill;
  M.prol = c->dxprol;
  M.flux = c->dxflux;
  M.nfill = (int)fill.size();
  M.nprol = (int)prol.size();
  M.nflux = (int)flux.size();
  return 0;
}

// Uploads a tile -> block table and the per-block first tile (nblocks + 1).
static int upload_tiles(pangu_ctx *c, const std::vector<int64_t> &tiles_per_block,
                        const int **tblock, const int **tfirst, int *ntiles) {
  std::vector<int> first(tiles_per_block.size() + 1, 0), owner;
  int64_t n = 0;
  for (size_t b = 0; b < tiles_per_block.size(); ++b) {
    first[b] = (int)n;
    for (int64_t t = 0; t < tiles_per_block[b]; ++t) owner.push_back((int)b);
    n += tiles_per_block[b];
  }
  first[tiles_per_block.size()] = (int)n;
  if (n > (int64_t)0x7fffff00) return fail(c, PANGU_EINVAL, "launch exceeds 2^31 tiles");
  int *dt = nullptr, *df = nullptr;
  if (upload(c, &dt, owner.data(), owner.size())) return PANGU_EDEVICE;
  c->tab_bufs.push_back(dt);
  if (upload(c, &df, first.data(), first.size())) return PANGU_EDEVICE;
  c->tab_bufs.push_back(df);
  *tblock = dt;
  *tfirst = df;
  *ntiles = (int)n;
  return 0;
}

// Host replica of the device block_range: tiles of one block for a range.
static int64_t range_tiles(const HostBlock &H, const int rlo[3], const int rhi[3], int entire,
                           int tile) {
  int cnt[3];
  for (int d = 0; d < 3; ++d) {
    int l, h;
    if (entire) {
      l = 0;
      h = H.n[d] - 1;
    } else if (H.n[d] > 1) {
      l = H.s[d] + rlo[d];
      h = H.e[d] + rhi[d];
    } else {
      l = H.s[d];
      h = H.e[d];
    }
    cnt[d] = h - l + 1;
    if (cnt[d] <= 0) return 0;
  }
  const int64_t plane = (int64_t)cnt[0] * cnt[1];
  return (int64_t)cnt[2] * ((plane + tile - 1) / tile);
}

static int make_range(pangu_ctx *c, Range3 &R, int l0, int l1, int l2, int h0, int h1, int h2,
                      int entire, int need_dim, int tile = 256) {
  R.rlo[0] = l0;
  R.rlo[1] = l1;
  R.rlo[2] = l2;
  R.rhi[0] = h0;
  R.rhi[1] = h1;
  R.rhi[2] = h2;
  R.entire = entire;
  R.order = 0;
  R.step = 0;
  R.tile = tile;
  std::vector<int64_t> tiles(c->nblocks);
  for (int b = 0; b < c->nblocks; ++b)
    tiles[b] = (need_dim >= 0 && c->hb[b].n[need_dim] == 1)
                   ? 0
                   : range_tiles(c->hb[b], R.rlo, R.rhi, entire, tile);
  // one block shape: arithmetic decoding
  R.uni = 0;
  bool same = true;
  for (int b = 1; b < c->nblocks && same; ++b)
    for (int d = 0; d < 3; ++d)
      if (c->hb[b].n[d] != c->hb[0].n[d] || c->hb[b].s[d] != c->hb[0].s[d] || c->hb[b].e[d] != c->hb[0].e[d])
        same = false;
  if (same && tiles[0] > 0) {
    const HostBlock &H = c->hb[0];
    for (int d = 0; d < 3; ++d) {
      int l, h;
      if (entire) {
        l = 0;
        h = H.n[d] - 1;
      } else if (H.n[d] > 1) {
        l = H.s[d] + R.rlo[d];
        h = H.e[d] + R.rhi[d];
      } else {
        l = H.s[d];
        h = H.e[d];
      }
      R.ulo[d] = l;
      R.ucnt[d] = h - l + 1;
    }
    R.utpp = (int)(((int64_t)R.ucnt[0] * R.ucnt[1] + tile - 1) / tile);
    R.uni = 1;
  }
  return upload_tiles(c, tiles, &R.tblock, &R.tfirst, &R.ntiles);
}

// Cell range of the shared-reconstruction X1 sweep: per plane, waves of 64
// lanes over the row-major cells (i = s-1 .. e+1) overlapping by one cell,
// four waves per tile.
static int make_range_x1s(pangu_ctx *c, Range3 &R) {
  R.rlo[0] = R.rlo[1] = R.rlo[2] = -1;
  R.rhi[0] = R.rhi[1] = R.rhi[2] = 1;
  R.entire = 0;
  R.order = 0;
  R.uni = 0;
  R.step = 0;
  R.tile = 256;
  std::vector<int64_t> tiles(c->nblocks, 0);
  for (int b = 0; b < c->nblocks; ++b) {
    const HostBlock &H = c->hb[b];
    if (H.n[0] == 1) continue;
    int cnt[3];
    for (int d = 0; d < 3; ++d)
      cnt[d] = H.n[d] > 1 ? (H.e[d] + 1) - (H.s[d] - 1) + 1 : 1;
    const int64_t plane = (int64_t)cnt[0] * cnt[1];
    const int64_t wpp = (plane - 1 + 62) / 63;
    tiles[b] = (int64_t)cnt[2] * ((wpp + 3) / 4);
  }
  return upload_tiles(c, tiles, &R.tblock, &R.tfirst, &R.ntiles);
}

// Marching range of the sweep along dir (2 or 3) over the face range F
// (decode_march): per block, ceil(count along dir / step) chunks of
// ceil(plane / 256) tiles, the plane spanned by x1 and the remaining
// direction.
static int make_range_march(pangu_ctx *c, Range3 &R, const Range3 &F, int dir, int step) {
  R = F;
  R.step = step;
  R.order = 0;
  const int D = dir - 1, E = (dir == 2) ? 2 : 1;
  std::vector<int64_t> tiles(c->nblocks, 0);
  for (int b = 0; b < c->nblocks; ++b) {
    const HostBlock &H = c->hb[b];
    if (H.n[D] == 1) continue;
    int cnt[3];
    bool empty = false;
    for (int d = 0; d < 3; ++d) {
      int l, h;
      if (F.entire) {
        l = 0;
        h = H.n[d] - 1;
      } else if (H.n[d] > 1) {
        l = H.s[d] + F.rlo[d];
        h = H.e[d] + F.rhi[d];
      } else {
        l = H.s[d];
        h = H.e[d];
      }
      cnt[d] = h - l + 1;
      if (cnt[d] <= 0) empty = true;
    }
    if (empty) continue;
    const int64_t plane = (int64_t)cnt[0] * cnt[E];
    tiles[b] = (int64_t)((cnt[D] + step - 1) / step) * ((plane + 255) / 256);
  }
  return upload_tiles(c, tiles, &R.tblock, &R.tfirst, &R.ntiles);
}

static int make_items(pangu_ctx *c, Items &I, const std::vector<int64_t> &items) {
  std::vector<int64_t> tiles(items.size());
  for (size_t b = 0; b < items.size(); ++b) tiles[b] = (items[b] + 255) / 256;
  return upload_tiles(c, tiles, &I.tblock, &I.tfirst, &I.ntiles);
}

// Launch tables of every range the kernels use.  Face ranges follow the
// three sweeps (laxf.cc:78-83,184-189,290-295): the normal index runs over
// s..e+1 and transverse indices include one ghost layer on active dims,
// except HLL's X3 sweep, which stays on the interior transversely
// (hll.cc:295-296).  CT covers s..e+1 on active dims.
static int build_tables(pangu_ctx *c) {
  LaunchCtx &L = c->L;
  // interior, entire, the three face ranges and CT; each in tiles of 256
  // cells and again of 128 and 64
  const bool hll = c->cfg.riemann == PANGU_RIEMANN_HLL;
  const int spec[6][8] = {{0, 0, 0, 0, 0, 0, 0, -1},  {0, 0, 0, 0, 0, 0, 1, -1},
                          {0, -1, -1, 1, 1, 1, 0, 0}, {-1, 0, -1, 1, 1, 1, 0, 1},
                          {hll ? 0 : -1, hll ? 0 : -1, 0, hll ? 0 : 1, hll ? 0 : 1, 1, 0, 2},
                          {0, 0, 0, 1, 1, 1, 0, -1}};
  Range3 *const base[6] = {&L.interior, &L.entire, &L.face[0], &L.face[1], &L.face[2], &L.ct};
  int r = 0;
  for (int id = 0; id < 6 && !r; ++id) {
    const int *q = spec[id];
    r = make_range(c, *base[id], q[0], q[1], q[2], q[3], q[4], q[5], q[6], q[7], 256);
    for (int z = 0; z < 2 && !r; ++z)
      r = make_range(c, L.small[id][z], q[0], q[1], q[2], q[3], q[4], q[5], q[6], q[7], z ? 64 : 128);
  }
  if (!r) r = make_range_x1s(c, L.x1cells);
  if (!r) r = make_range(c, L.fixx1, 0, 0, 0, 0, 0, 0, 1, -1, 248);
  std::vector<int64_t> ghosts(c->nblocks), planes(c->nblocks);
  for (int b = 0; b < c->nblocks; ++b) {
    const HostBlock &H = c->hb[b];
    const int64_t interior = (int64_t)(H.e[0] - H.s[0] + 1) * (H.e[1] - H.s[1] + 1) * (H.e[2] - H.s[2] + 1);
    ghosts[b] = H.cells - interior;
    planes[b] = H.e[2] - H.s[2] + 1;
  }
  if (!r) r = make_items(c, L.ghosts, ghosts);
  for (int d = 0; d < 3 && !r; ++d) {
    std::vector<int64_t> slab(c->nblocks, 0);
    for (int b = 0; b < c->nblocks; ++b) {
      const HostBlock &H = c->hb[b];
      const pangu_config &cfg = c->mcfg[c->bmesh[b]];
      if (H.n[d] == 1) continue;
      const bool lo = H.lx[d] == 0 && cfg.bc[d][0] != PANGU_BC_PERIODIC;
      const bool hi = H.lx[d] == cfg.nb[d] - 1 && cfg.bc[d][1] != PANGU_BC_PERIODIC;
      if (!lo && !hi) continue;
      const int other = d == 0 ? H.n[1] * H.n[2] : (d == 1 ? H.n[0] * H.n[2] : H.n[0] * H.n[1]);
      slab[b] = 2 * (int64_t)H.s[d] * other;
    }
    r = make_items(c, L.bcs[d], slab);
  }
  if (r) return r;
  // interior planes: one entry per (block, k) plane (dt reduction, FixRecovery)
  const int *pb = nullptr, *pf = nullptr;
  int np = 0;
  if (int e = upload_tiles(c, planes, &pb, &pf, &np)) return e;
  c->nplanes = np;
  L.k.pblock = pb;
  L.k.pfirst = pf;
  L.k.nplanes = np;
  // dt reduction items: every interior plane in chunks of 256 * dtper cells,
  // so that a large single-block plane is scanned by many workgroups
  int64_t maxplane = 1;
  for (int b = 0; b < c->nblocks; ++b) {
    const HostBlock &H = c->hb[b];
    maxplane = std::max(maxplane, (int64_t)(H.e[0] - H.s[0] + 1) * (H.e[1] - H.s[1] + 1));
  }
  L.k.dtper = dt_per_thread(maxplane);
  const int64_t chunk = 256 * (int64_t)L.k.dtper;
  std::vector<int64_t> dtitems(c->nblocks);
  for (int b = 0; b < c->nblocks; ++b) {
    const HostBlock &H = c->hb[b];
    const int64_t plane = (int64_t)(H.e[0] - H.s[0] + 1) * (H.e[1] - H.s[1] + 1);
    dtitems[b] = planes[b] * ((plane + chunk - 1) / chunk);
  }
  const int *db = nullptr, *dfi = nullptr;
  int nd = 0;
  if (int e = upload_tiles(c, dtitems, &db, &dfi, &nd)) return e;
  c->ndt = nd;
  L.k.dtblock = db;
  L.k.dtfirst = dfi;
  L.k.ndt = nd;
  int *dm = nullptr;
  if (upload(c, &dm, c->mfirst.data(), c->mfirst.size())) return PANGU_EDEVICE;
  c->tab_bufs.push_back(dm);
  L.k.mfirst = dm;
  return 0;
}

namespace pangu {
void set_kernel_clock(pangu_ctx *c, KernelClock *clock) {
  if (c) c->L.clock = clock;
}
}  // namespace pangu

extern "C" {

int pangu_create_batch(const pangu_config *cfgs, int nmesh, const pangu_block *blocks,
                       const int *block_mesh, int nblocks, int device, pangu_ctx **out) {
  if (!out) return PANGU_EINVAL;
  *out = nullptr;
  std::string msg;
  if (!validate_batch(cfgs, nmesh, blocks, block_mesh, nblocks, msg)) {
    fprintf(stderr, "pangu_create: %s\n", msg.c_str());
    return PANGU_EINVAL;
  }
  RefineTables rt;
  {
    std::vector<HostBlock> hb;
    std::vector<int> bm(nblocks);
    for (int b = 0; b < nblocks; ++b) {
      hb.push_back(make_host_block(blocks[b]));
      bm[b] = block_mesh ? block_mesh[b] : 0;
    }
    msg = plan_multilevel(hb, bm, std::vector<pangu_config>(cfgs, cfgs + nmesh), rt);
    if (!msg.empty()) {
      fprintf(stderr, "pangu_create: %s\n", msg.c_str());
      return PANGU_EINVAL;
    }
  }
  pangu_ctx *c = new pangu_ctx();
  const pangu_config *cfg = &cfgs[0];
  c->cfg = *cfg;
  c->mcfg.assign(cfgs, cfgs + nmesh);
  c->nmesh = nmesh;
  c->device = device;
  c->nblocks = nblocks;
  c->bmesh.resize(nblocks);
  c->mfirst.assign(nmesh + 1, nblocks);
  for (int b = nblocks - 1; b >= 0; --b) {
    c->bmesh[b] = block_mesh ? block_mesh[b] : 0;
    c->mfirst[c->bmesh[b]] = b;
  }
  if (hipSetDevice(device) != hipSuccess) {
    delete c;
    return PANGU_EDEVICE;
  }
  c->ndim = 1;
  for (int b = 0; b < nblocks; ++b) {
    const int nd = (blocks[b].nx[2] > 1) ? 3 : (blocks[b].nx[1] > 1 ? 2 : 1);
    if (nd > c->ndim) c->ndim = nd;
  }
  // packed component layout (package_registration.cc:84-96)
  c->np = 5;
  if (cfg->enable_B) {
    c->iBX = c->np;
    c->np += 3;
  }
  c->iENT = c->np++;
  if (cfg->enable_heating) c->iKEL = c->np++;
  c->nc = c->np;
  for (int b = 0; b < nblocks; ++b) c->hb.push_back(make_host_block(blocks[b]));
  int r = build_geometry(c, blocks);
  if (!r) r = build_neighbours(c);
  if (!r) r = build_multilevel(c, rt);
  if (!r) r = build_tables(c);
  if (!r) r = upload(c, &c->dred, std::vector<double>((size_t)c->ndt * 3, 0.0).data(),
                     (size_t)c->ndt * 3);
  if (!r) r = upload(c, &c->dcluster, std::vector<int>((size_t)c->nplanes, 0).data(),
                     (size_t)c->nplanes);
  if (!r) r = upload(c, &c->dqueue, std::vector<unsigned>(2, 0u).data(), 2);
  if (!r) c->L.queue = c->dqueue;
  if (!r) r = upload(c, &c->dest, std::vector<double>(nmesh, 0.0).data(), nmesh);
  if (!r) {
    // interior tiles are grouped by block and therefore by mesh
    c->diag_mtile.assign(nmesh + 1, 0);
    const int zero[3] = {0, 0, 0};
    int64_t tot = 0;
    for (int m = 0; m < nmesh; ++m) {
      c->diag_mtile[m] = (int)tot;
      for (int b = c->mfirst[m]; b < c->mfirst[m + 1]; ++b) tot += range_tiles(c->hb[b], zero, zero, 0, 256);
    }
    c->diag_mtile[nmesh] = (int)tot;
    if (tot != c->L.interior.ntiles) r = fail(c, PANGU_EINVAL, "diagnostics: tile count mismatch");
  }
  if (r) {
    fprintf(stderr, "pangu_create: %s\n", c->err.c_str());
    pangu_destroy(c);
    return r;
  }
  {
    // address of this device's static diagnostics workspace, looked up once
    // per device with the device current; a failure here only disables
    // pangu_diagnostics_noalloc
    const hipError_t e = diag_scratch_address(device, &c->dscratch);
    if (e != hipSuccess) {
      c->dscratch = nullptr;
      c->scratch_err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
    }
  }
  if (hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming) != hipSuccess) {
    c->err = "pangu_create: side stream or events";
    fprintf(stderr, "%s\n", c->err.c_str());
    pangu_destroy(c);
    return PANGU_EDEVICE;
  }
  KCtx &k = c->L.k;
  k.blk = c->dblk;
  k.geo = c->dgeo;
  k.nblocks = nblocks;
  k.nmesh = nmesh;
  k.np = c->np;
  k.nc = c->nc;
  k.iBX = c->iBX;
  k.iENT = c->iENT;
  k.iKEL = c->iKEL;
  k.ndim = c->ndim;
  Phys &P = k.ph;
  P.gamma = cfg->gamma;
  P.cfl = cfg->cfl;
  P.density_floor = cfg->density_floor;
  P.energy_floor = cfg->energy_floor;
  P.density_floor_pow = cfg->density_floor_pow;
  P.energy_floor_pow = cfg->energy_floor_pow;
  P.sigma_max = cfg->sigma_max;
  P.lorentz_max = cfg->lorentz_max;
  P.heat_model = cfg->heat_model;
  P.fel_constant = cfg->fel_constant;
  P.gamma_e = cfg->gamma_e;
  P.gamma_p = cfg->gamma_p;
  P.ratio_min = cfg->ratio_min;
  P.ratio_max = cfg->ratio_max;
  P.limit_kel = cfg->limit_kel;
  P.suppress_highb_heat = cfg->suppress_highb_heat;
  P.enforce_positive_dissipation = cfg->enforce_positive_dissipation;
  P.metric = cfg->metric;
  P.excise = cfg->excise;
  P.r_excise = cfg->r_excise;
  P.dexcise = cfg->dexcise;
  P.pexcise = cfg->pexcise;
  P.a = cfg->a;
  c->L.limiter = cfg->limiter;
  c->L.riemann = cfg->riemann;
  c->L.metric = cfg->metric;
  c->L.recovery = cfg->recovery;
  *out = c;
  return PANGU_OK;
}

int pangu_create(const pangu_config *cfg, const pangu_block *blocks, int nblocks, int device,
                 pangu_ctx **out) {
  return pangu_create_batch(cfg, 1, blocks, nullptr, nblocks, device, out);
}

int pangu_destroy(pangu_ctx *c) {
  if (!c) return PANGU_OK;
  (void)hipSetDevice(c->device);
  for (double *p : c->geo_bufs) (void)hipFree(p);
  for (int *p : c->tab_bufs) (void)hipFree(p);
  for (auto &ct : c->conts) {
    (void)hipFree(ct.dprim);
    (void)hipFree(ct.dcons);
    for (int d = 0; d < 3; ++d) (void)hipFree(ct.dflux[d]);
  }
  void *ptrs[] = {c->dblk,    c->dgeo,  c->dnbr,     c->dalfven, c->demf,  c->dflag,
                  c->dflag_rec, c->dmask, c->dstatus, c->diters,  c->droot, c->dred,  c->dcluster,
                  c->dqueue,  c->dest,    c->ddiag_part, c->ddiag_mtile,
                  c->ddiag_shell, c->dxfill, c->dxprol, c->dxflux};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  if (c->side) {
    (void)hipStreamSynchronize(c->side);
    (void)hipStreamDestroy(c->side);
  }
  if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
  if (c->ev_join) (void)hipEventDestroy(c->ev_join);
  delete c;
  return PANGU_OK;
}

int pangu_set_tile_order(pangu_ctx *c, int fam, int order) {
  if (!c) return PANGU_EINVAL;
  if (fam < 0 || fam >= PANGU_KF_COUNT || (order != PANGU_ORDER_PLANE && order != PANGU_ORDER_COLUMN))
    return fail(c, PANGU_EINVAL, "tile order: unknown kernel family or order");
  c->L.order[fam] = order;
  return PANGU_OK;
}

int pangu_set_fuse_ct(pangu_ctx *c, int fuse) {
  if (!c) return PANGU_EINVAL;
  c->L.fuse_ct = fuse != 0;
  return PANGU_OK;
}

int pangu_set_fuse_fix_x1(pangu_ctx *c, int on) {
  if (!c) return PANGU_EINVAL;
  c->L.fuse_fix_x1 = on != 0;
  return PANGU_OK;
}

int pangu_set_fuse_update_recovery(pangu_ctx *c, int on) {
  if (!c) return PANGU_EINVAL;
  c->L.fuse_update_recovery = on != 0;
  return PANGU_OK;
}

int pangu_set_overlap_exchange(pangu_ctx *c, int on) {
  if (!c) return PANGU_EINVAL;
  if (on < -1 || on > 1) return fail(c, PANGU_EINVAL, "overlap_exchange: -1, 0 or 1");
  c->L.overlap_exchange = on;
  return PANGU_OK;
}

int pangu_set_split_source(pangu_ctx *c, int split) {
  if (!c) return PANGU_EINVAL;
  c->L.split_source = split != 0;
  return PANGU_OK;
}

int pangu_set_c2p_guess(pangu_ctx *c, pangu_mlp *m, double delta) {
  if (!c) return PANGU_EINVAL;
  if (m && !(delta > 0.0 && delta < 1.0)) return fail(c, PANGU_EINVAL, "delta must be in (0, 1)");
  if (m && mlp_kind(m) == 0 && c->cfg.recovery != PANGU_RECOVERY_ROBUSTER)
    return fail(c, PANGU_EINVAL, "a Kastaun-bracket model applies to the robuster (Kastaun) recovery");
  if (m && mlp_kind(m) == 1 && c->cfg.recovery != PANGU_RECOVERY_ONEDVSQ)
    return fail(c, PANGU_EINVAL, "a 1Dvsq-start model applies to the onedvsq recovery");
  MlpFrag W;
  if (m && !mlp_view(m, W)) return fail(c, PANGU_ESTATE, "learned recovery model has no weights");
  c->mlp = m;
  c->guess_delta = m ? delta : 0.0;
  return PANGU_OK;
}

int pangu_set_c2p_gate(pangu_ctx *c, int min_prev_iters) {
  if (!c) return PANGU_EINVAL;
  if (min_prev_iters < 0) return fail(c, PANGU_EINVAL, "gate must be >= 0");
  c->guess_gate = min_prev_iters;
  return PANGU_OK;
}


int pangu_set_pair_flux(pangu_ctx *c, int dir_mask) {
  if (!c) return PANGU_EINVAL;
  if (dir_mask < 0 || dir_mask > 7) return fail(c, PANGU_EINVAL, "pair flux: mask of directions 0..7");
  c->L.pair_flux = dir_mask;
  return PANGU_OK;
}

int pangu_set_tile_size(pangu_ctx *c, int fam, int cells) {
  if (!c || fam < 0 || fam >= PANGU_KF_COUNT) return PANGU_EINVAL;
  if (cells != 64 && cells != 128 && cells != 256)
    return fail(c, PANGU_EINVAL, "tile size: 64, 128 or 256 cells");
  c->L.tile[fam] = cells == 256 ? 0 : cells;
  return PANGU_OK;
}

int pangu_get_tile_size(const pangu_ctx *c, int fam) {
  if (!c || fam < 0 || fam >= PANGU_KF_COUNT) return PANGU_EINVAL;
  return c->L.tile[fam] ? c->L.tile[fam] : 256;
}

int pangu_set_persistent_recovery(pangu_ctx *c, int on) {
  if (!c) return PANGU_EINVAL;
  c->L.persist_recovery = on != 0;
  return PANGU_OK;
}

int pangu_set_flux_march(pangu_ctx *c, int dir, int depth) {
  if (!c) return PANGU_EINVAL;
  if (dir != 2 && dir != 3) return fail(c, PANGU_EINVAL, "flux march: direction 2 or 3");
  if (depth < 0 || depth > 64) return fail(c, PANGU_EINVAL, "flux march: depth 0..64");
  LaunchCtx &L = c->L;
  if (depth > 0 && (L.march_depth[dir - 2] != depth || L.march[dir - 2].ntiles == 0)) {
    // the face ranges are built on creation; the tables of a previous depth
    // stay with the context until it is destroyed
    Range3 R{};
    if (int r = make_range_march(c, R, L.face[dir - 1], dir, depth)) return fail(c, r, "flux march: tables");
    L.march[dir - 2] = R;
  }
  L.march_depth[dir - 2] = depth;
  return PANGU_OK;
}

int pangu_set_share_reconstruction(pangu_ctx *c, int share) {
  if (!c) return PANGU_EINVAL;
  c->L.share_x1 = share != 0;
  return PANGU_OK;
}

int pangu_get_tile_order(const pangu_ctx *c, int fam) {
  if (!c || fam < 0 || fam >= PANGU_KF_COUNT) return PANGU_EINVAL;
  return c->L.order[fam];
}

const char *pangu_last_error(const pangu_ctx *c) { return c ? c->err.c_str() : "null context"; }

int pangu_geometry_entry(const pangu_config *cfg, int loc, const double x[4], double *out) {
  if (!cfg || !x || !out || loc < 0 || loc > 3) return PANGU_EINVAL;
  if (loc == 0)
    geom::fill_centre(*cfg, x, out, 1);
  else
    geom::fill_face(*cfg, loc - 1, x, out, 1);
  return PANGU_OK;
}

int pangu_register_container(pangu_ctx *c, const pangu_container *src, int *handle) {
  if (!c || !src || !handle || !src->prim || !src->cons) return fail(c, PANGU_EINVAL, "null container");
  HIPCHK(c, hipSetDevice(c->device));
  DevContainer d;
  d.hprim.assign(src->prim, src->prim + (size_t)c->nblocks * c->np);
  d.hcons.assign(src->cons, src->cons + (size_t)c->nblocks * c->nc);
  if (upload(c, &d.dprim, d.hprim.data(), d.hprim.size())) return PANGU_EDEVICE;
  if (upload(c, &d.dcons, d.hcons.data(), d.hcons.size())) return PANGU_EDEVICE;
  for (int k = 0; k < 3; ++k) {
    if (src->flux[k]) {
      d.hflux[k].assign(src->flux[k], src->flux[k] + (size_t)c->nblocks * c->nc);
      if (upload(c, &d.dflux[k], d.hflux[k].data(), d.hflux[k].size())) return PANGU_EDEVICE;
    } else if (k < c->ndim) {
      d.dflux[k] = nullptr;
    }
  }
  *handle = (int)c->conts.size();
  c->conts.push_back(d);
  return PANGU_OK;
}

int pangu_register_shared(pangu_ctx *c, const pangu_shared *s) {
  if (!c || !s || !s->alfven) return fail(c, PANGU_EINVAL, "shared fields need alfven");
  HIPCHK(c, hipSetDevice(c->device));
  const size_t nb = (size_t)c->nblocks;
  // a second registration replaces the tables of the first
  HIPCHK(c, hipDeviceSynchronize());
  for (void **p : {(void **)&c->dalfven, (void **)&c->demf, (void **)&c->dflag, (void **)&c->dflag_rec,
                   (void **)&c->dmask, (void **)&c->dstatus, (void **)&c->diters, (void **)&c->droot}) {
    if (*p) (void)hipFree(*p);
    *p = nullptr;
  }
  c->shared_ok = false;
  if (upload(c, &c->dalfven, s->alfven, nb * 3)) return PANGU_EDEVICE;
  if (s->emf && upload(c, &c->demf, s->emf, nb * 3)) return PANGU_EDEVICE;
  if (s->flag && upload(c, &c->dflag, s->flag, nb)) return PANGU_EDEVICE;
  if (s->flag_recovery && upload(c, &c->dflag_rec, s->flag_recovery, nb)) return PANGU_EDEVICE;
  if (s->mask && upload(c, &c->dmask, s->mask, nb)) return PANGU_EDEVICE;
  if (s->c2p_status && upload(c, &c->dstatus, s->c2p_status, nb)) return PANGU_EDEVICE;
  if (s->c2p_iters && upload(c, &c->diters, s->c2p_iters, nb)) return PANGU_EDEVICE;
  if (s->c2p_root && upload(c, &c->droot, s->c2p_root, nb)) return PANGU_EDEVICE;
  if (c->cfg.enable_B && !c->demf) return fail(c, PANGU_EINVAL, "enable_B needs emf tables");
  if (c->cfg.recovery != PANGU_RECOVERY_ROBUSTER && !c->dflag)
    return fail(c, PANGU_EINVAL, "chainer/onedvsq recovery needs the flag field");
  c->shared_ok = true;
  return PANGU_OK;
}

static int check_md(pangu_ctx *c, int md) {
  if (!c) return PANGU_EINVAL;
  if (md < 0 || md >= (int)c->conts.size()) return fail(c, PANGU_ESTATE, "unknown container handle");
  if (!c->shared_ok) return fail(c, PANGU_ESTATE, "pangu_register_shared not called");
  return 0;
}

int pangu_mlp_harvest(pangu_mlp *m, pangu_ctx *c, int md, int stride, int64_t *kept) {
  if (int r = check_md(c, md)) return r;
  if (!m || stride < 1) return fail(c, PANGU_EINVAL, "harvest: null model or stride < 1");
  if (c->device != mlp_device(m)) return fail(c, PANGU_EINVAL, "harvest: model and context on different devices");
  HIPCHK(c, hipSetDevice(c->device));
  // the container's fields are written on the caller's streams
  HIPCHK(c, hipDeviceSynchronize());
  if (int r = mlp_harvest(m, c->L, c->conts[md].kf(), stride, kept))
    return fail(c, r, std::string("harvest: ") + pangu_mlp_last_error(m));
  return PANGU_OK;
}

#define STREAM(s) ((hipStream_t)(s))
#define LCHK(c, x)                                                                  \
  do {                                                                              \
    if ((x) != 0) return fail(c, PANGU_EDEVICE, std::string("kernel launch failed: ") + \
                                                   hipGetErrorString(hipGetLastError())); \
  } while (0)

int pangu_fix_primitive(pangu_ctx *c, int md, void *stream) {
  if (int r = check_md(c, md)) return r;
  LCHK(c, launch_fix_cons(c->L, c->conts[md].kf(), c->dmask, true, false, STREAM(stream)));
  return PANGU_OK;
}

int pangu_calculate_conservative(pangu_ctx *c, int md, void *stream) {
  if (int r = check_md(c, md)) return r;
  LCHK(c, launch_fix_cons(c->L, c->conts[md].kf(), nullptr, false, true, STREAM(stream)));
  return PANGU_OK;
}

// Entry points that read or write face fluxes need the container's flux
// tables on every active direction.
static int need_fluxes(pangu_ctx *c, int md) {
  for (int d = 0; d < c->ndim; ++d)
    if (!c->conts[md].dflux[d]) return fail(c, PANGU_EINVAL, "container has no flux tables");
  return 0;
}

int pangu_calculate_fluxes(pangu_ctx *c, int md, void *stream) {
  if (int r = check_md(c, md)) return r;
  if (int r = need_fluxes(c, md)) return r;
  LCHK(c, launch_fluxes(c->L, c->conts[md].kf(), c->dalfven, STREAM(stream)));
  return PANGU_OK;
}

int pangu_constrained_transport(pangu_ctx *c, int md, void *stream) {
  if (int r = check_md(c, md)) return r;
  if (int r = need_fluxes(c, md)) return r;
  if (!c->cfg.enable_B) return PANGU_OK;
  LCHK(c, launch_ct(c->L, c->conts[md].kf(), c->demf, STREAM(stream)));
  return PANGU_OK;
}

int pangu_flux_divergence(pangu_ctx *c, int md_in, int md_dudt, void *stream) {
  if (int r = check_md(c, md_in)) return r;
  if (int r = need_fluxes(c, md_in)) return r;
  if (int r = check_md(c, md_dudt)) return r;
  const KFields f = c->conts[md_in].kf();
  LCHK(c, launch_div_update(c->L, 1, f, f, f, c->conts[md_dudt].dcons, false, nullptr, 0.0,
                            0.0, STREAM(stream)));
  return PANGU_OK;
}

int pangu_update_independent(pangu_ctx *c, int md_init, int md_dudt, double beta_dt, int md_out,
                             void *stream) {
  if (int r = check_md(c, md_init)) return r;
  if (int r = check_md(c, md_dudt)) return r;
  if (int r = check_md(c, md_out)) return r;
  LCHK(c, launch_div_update(c->L, 2, c->conts[md_init].kf(), c->conts[md_init].kf(),
                            c->conts[md_out].kf(), c->conts[md_dudt].dcons, false, nullptr, 0.0,
                            beta_dt, STREAM(stream)));
  return PANGU_OK;
}

int pangu_add_geometric_source(pangu_ctx *c, int md, double beta_dt, void *stream) {
  if (int r = check_md(c, md)) return r;
  const KFields f = c->conts[md].kf();
  LCHK(c, launch_div_update(c->L, 3, f, f, f, nullptr, false, nullptr, 0.0, beta_dt,
                            STREAM(stream)));
  return PANGU_OK;
}

static SharedTabs shared_tabs(pangu_ctx *c) {
  SharedTabs S;
  S.flag = c->dflag;
  S.flag_rec = c->dflag_rec;
  S.status = c->dstatus;
  S.iters = c->diters;
  S.root = c->droot;
  S.delta = c->guess_delta;
  S.gate = c->guess_gate;
  return S;
}

// The learned Kastaun bracket's weights when a model is attached (and the
// recovery is Kastaun), else null.
static const MlpFrag *learned(pangu_ctx *c, MlpFrag &W) {
  if (!c->mlp) return nullptr;
  const int kind = mlp_kind(c->mlp);
  if (!((kind == 0 && c->cfg.recovery == PANGU_RECOVERY_ROBUSTER) ||
        (kind == 1 && c->cfg.recovery == PANGU_RECOVERY_ONEDVSQ)))
    return nullptr;
  return mlp_view(c->mlp, W) ? &W : nullptr;
}

int pangu_recovery(pangu_ctx *c, int md, void *stream) {
  if (int r = check_md(c, md)) return r;
  MlpFrag W;
  LCHK(c, launch_recover(c->L, c->conts[md].kf(), shared_tabs(c), c->cfg.recovery, false,
                         learned(c, W), STREAM(stream)));
  return PANGU_OK;
}

int pangu_apply_electron_heating(pangu_ctx *c, int md, void *stream) {
  if (int r = check_md(c, md)) return r;
  if (!c->cfg.enable_heating) return PANGU_OK;
  LCHK(c, launch_recover(c->L, c->conts[md].kf(), shared_tabs(c), -1, true, nullptr, STREAM(stream)));
  return PANGU_OK;
}

int pangu_fix_recovery(pangu_ctx *c, int md, void *stream) {
  if (int r = check_md(c, md)) return r;
  if (!c->dflag) return PANGU_OK;
  LCHK(c, launch_fix_recovery(c->L, c->conts[md].kf(), c->dflag, c->dcluster, STREAM(stream)));
  return PANGU_OK;
}

int pangu_estimate_timestep(pangu_ctx *c, int md, double *dt_out, void *stream) {
  if (int r = check_md(c, md)) return r;
  if (!dt_out) return fail(c, PANGU_EINVAL, "dt_out is null");
  LCHK(c, launch_dt(c->L, c->dalfven, c->dred, nullptr, 1.0, c->dest, STREAM(stream)));
  HIPCHK(c, hipMemcpyAsync(dt_out, c->dest, sizeof(double) * c->nmesh, hipMemcpyDeviceToHost,
                           STREAM(stream)));
  HIPCHK(c, hipStreamSynchronize(STREAM(stream)));
  return PANGU_OK;
}

static int exchange(pangu_ctx *c, const KFields &f, hipStream_t s) {
  return c->L.ml.on ? launch_exchange_ml(c->L, f, c->dflag, s) : launch_exchange(c->L, f, c->dflag, c->dnbr, s);
}

int pangu_exchange_ghosts(pangu_ctx *c, int md, void *stream) {
  if (int r = check_md(c, md)) return r;
  LCHK(c, exchange(c, c->conts[md].kf(), STREAM(stream)));
  return PANGU_OK;
}

int pangu_flux_correction(pangu_ctx *c, int md, void *stream) {
  if (int r = check_md(c, md)) return r;
  if (int r = need_fluxes(c, md)) return r;
  LCHK(c, launch_flux_correction(c->L, c->conts[md].kf(), STREAM(stream)));
  return PANGU_OK;
}

int pangu_apply_boundary_conditions(pangu_ctx *c, int md, void *stream) {
  if (int r = check_md(c, md)) return r;
  LCHK(c, launch_bc(c->L, c->conts[md].kf(), c->dflag, STREAM(stream)));
  return PANGU_OK;
}

int pangu_stage(pangu_ctx *c, int md_init, int md0, int md1, double beta, double *dt_state,
                double dt_factor, int last_stage, int exchange_after, void *stream) {
  if (int r = check_md(c, md_init)) return r;
  if (int r = check_md(c, md0)) return r;
  if (int r = check_md(c, md1)) return r;
  if (int r = need_fluxes(c, md0)) return r;
  if (!dt_state) return fail(c, PANGU_EINVAL, "dt_state is null");
  hipStream_t s = STREAM(stream);
  const KFields f0 = c->conts[md0].kf(), fi = c->conts[md_init].kf(), f1 = c->conts[md1].kf();
  // The exchange after the update copies interior P and flag of the output
  // container into its ghosts.  Nothing before the recovery writes that
  // interior P (the update writes U, FixPrimitive the input container's P)
  // or reads those ghosts or the flag, so with distinct input and output
  // containers it runs on the side stream from the start of the stage,
  // beside FixPrimitive and the flux sweeps, and joins before the recovery.
  // (Sequential when a kernel clock times the launches on one stream.)  By
  // default only on a refined mesh: elsewhere the cross-stream join costs
  // more than the overlap hides.
  const bool want = c->L.overlap_exchange > 0 || (c->L.overlap_exchange < 0 && c->L.ml.on);
  const bool overlap = exchange_after == 0 && md1 != md0 && want && c->side && !c->L.clock;
  if (overlap) {
    HIPCHK(c, hipEventRecord(c->ev_fork, s));
    HIPCHK(c, hipStreamWaitEvent(c->side, c->ev_fork, 0));
    LCHK(c, exchange(c, f1, c->side));
    HIPCHK(c, hipEventRecord(c->ev_join, c->side));
  }
  // FixPrimitive + conservative and the X1 sweep in one kernel (the plain
  // per-face X1 form only)
  const bool fuse_fx = c->L.fuse_fix_x1 && !c->L.share_x1 && !(c->L.pair_flux & 1);
  if (fuse_fx) {
    LCHK(c, launch_fix_flux_x1(c->L, f0, c->dmask, c->dalfven, s));
    LCHK(c, launch_fluxes_23(c->L, f0, c->dalfven, s));
  } else {
    LCHK(c, launch_fix_cons(c->L, f0, c->dmask, true, true, s));
    LCHK(c, launch_fluxes(c->L, f0, c->dalfven, s));
  }
  // With fuse_ct the update kernel forms the edge EMFs and the CT face B
  // fluxes itself, so no CT kernel runs: the flux arrays keep the Riemann B
  // fluxes and the EMF arrays their last values after a fused stage
  // (nothing after the update reads them).  Flux correction on a refined
  // mesh overwrites face fluxes, so there CT writes them first.
  const bool fuse_ct = c->cfg.enable_B && c->L.fuse_ct && !c->L.ml.on;
  if (c->cfg.enable_B && !fuse_ct) LCHK(c, launch_ct(c->L, f0, c->demf, s));
  if (c->L.ml.on) LCHK(c, launch_flux_correction(c->L, f0, s));
  MlpFrag W;
  const MlpFrag *Wl = learned(c, W);
  // update and recovery in one kernel: the stage's exchange (after the
  // update) then runs before it, which changes nothing (it touches neither
  // U nor the update's inputs)
  const bool hb = c->L.k.iBX >= 0, hk = c->L.k.iKEL >= 0;
  const bool fuse_ur = c->L.fuse_update_recovery && (hb ? fuse_ct : !hk) && !c->L.split_source &&
                       !c->L.persist_recovery && c->cfg.recovery >= 0 && c->cfg.recovery <= 2;
  if (fuse_ur) {
    if (overlap)
      HIPCHK(c, hipStreamWaitEvent(s, c->ev_join, 0));
    else if (exchange_after == 0)
      LCHK(c, exchange(c, f1, s));
    if (Wl) {
      // the learned kernel of the scheme, with the update run first
      UpdArgs u;
      u.F0 = f0;
      u.Finit = fi;
      u.dt_dev = dt_state;
      u.beta = beta;
      LCHK(c, launch_recover(c->L, f1, shared_tabs(c), c->cfg.recovery, c->cfg.enable_heating != 0, Wl, s,
                             &u));
    } else {
      LCHK(c, launch_update_recover(c->L, f0, fi, f1, shared_tabs(c), c->cfg.recovery,
                                    c->cfg.enable_heating != 0, fuse_ct, dt_state, beta, 0.0, s));
    }
  } else if (c->L.split_source && c->cfg.metric != PANGU_METRIC_MINKOWSKI) {
    LCHK(c, launch_div_update(c->L, 5, f0, fi, f1, nullptr, fuse_ct, dt_state, beta, 0.0, s));
    LCHK(c, launch_div_update(c->L, 3, f0, fi, f1, nullptr, false, dt_state, beta, 0.0, s));
  } else {
    LCHK(c, launch_div_update(c->L, 4, f0, fi, f1, nullptr, fuse_ct, dt_state, beta, 0.0, s));
  }
  if (!fuse_ur) {
    if (overlap)
      HIPCHK(c, hipStreamWaitEvent(s, c->ev_join, 0));
    else if (exchange_after == 0)
      LCHK(c, exchange(c, f1, s));
    LCHK(c, launch_recover(c->L, f1, shared_tabs(c), c->cfg.recovery, c->cfg.enable_heating != 0, Wl,
                           s));
  }
  if (c->cfg.recovery != PANGU_RECOVERY_ROBUSTER)
    LCHK(c, launch_fix_recovery(c->L, f1, c->dflag, c->dcluster, s));
  if (exchange_after != 0) LCHK(c, exchange(c, f1, s));
  LCHK(c, launch_bc(c->L, f1, c->dflag, s));
  if (last_stage) LCHK(c, launch_dt(c->L, c->dalfven, c->dred, dt_state, dt_factor, nullptr, s));
  return PANGU_OK;
}

}  // extern "C"

static int check_diag_args(pangu_ctx *c, int md, const int *shell, const double *out) {
  if (int r = check_md(c, md)) return r;
  if (!shell || !out) return fail(c, PANGU_EINVAL, "diagnostics: null shell or output array");
  for (int m = 0; m < c->nmesh; ++m) {
    const HostBlock &H = c->hb[c->mfirst[m]];
    const int nx1 = c->mcfg[m].nb[0] * (H.e[0] - H.s[0] + 1);
    if (shell[m] < -1 || shell[m] >= nx1)
      return fail(c, PANGU_EINVAL, "diagnostics: shell index outside the mesh's x1 cells");
    if (shell[m] >= 0 && c->L.ml.on)
      return fail(c, PANGU_EINVAL,
                  "diagnostics: shell integrals need a single-level mesh (pass shell -1 on a refined one)");
  }
  return 0;
}

int pangu_diagnostics(pangu_ctx *c, int md, const int *shell, double *out, void *stream) {
  if (int r = check_diag_args(c, md, shell, out)) return r;
  HIPCHK(c, hipSetDevice(c->device));
  if (!c->ddiag_part) {
    const int64_t tot = c->diag_mtile[c->nmesh];
    HIPCHK(c, hipMalloc((void **)&c->ddiag_part, (size_t)(tot > 0 ? tot : 1) * PANGU_NDIAG * sizeof(double)));
    if (int r = upload(c, &c->ddiag_mtile, c->diag_mtile.data(), c->diag_mtile.size())) return r;
    HIPCHK(c, hipMalloc((void **)&c->ddiag_shell, (size_t)c->nmesh * sizeof(int)));
  }
  hipStream_t s = STREAM(stream);
  HIPCHK(c, hipMemcpyAsync(c->ddiag_shell, shell, (size_t)c->nmesh * sizeof(int), hipMemcpyHostToDevice, s));
  LCHK(c, launch_diagnostics(c->L, c->conts[md].kf(), c->ddiag_shell, c->ddiag_part, c->ddiag_mtile,
                             c->nmesh, out, s));
  // the shell indices are read from pageable host memory: wait before returning
  HIPCHK(c, hipStreamSynchronize(s));
  return PANGU_OK;
}

int pangu_diagnostics_noalloc(pangu_ctx *c, int md, const int *shell, double *out, void *stream) {
  if (int r = check_diag_args(c, md, shell, out)) return r;
  if (!c->dscratch)
    return fail(c, PANGU_EDEVICE, "diagnostics_noalloc: no static workspace on this device (" +
                                      c->scratch_err + ")");
  // workspace: tiles x PANGU_NDIAG partials, then nmesh + 1 tile offsets and
  // nmesh shell indices (ints, rounded up to whole doubles)
  const int64_t tiles = c->diag_mtile[c->nmesh];
  const int64_t nint = 2 * (int64_t)c->nmesh + 1;
  const int64_t need = tiles * PANGU_NDIAG + (nint + 1) / 2;
  if (need > kDiagScratchDoubles)
    return fail(c, PANGU_EINVAL,
                "diagnostics_noalloc: this context needs " + std::to_string(need) +
                    " doubles of workspace (" + std::to_string(tiles) +
                    " interior tiles); the static workspace holds " +
                    std::to_string(kDiagScratchDoubles) + " (8 MiB): use pangu_diagnostics");
  HIPCHK(c, hipSetDevice(c->device));
  hipStream_t s = STREAM(stream);
  double *part = c->dscratch;
  int *mtile = (int *)(c->dscratch + tiles * PANGU_NDIAG);
  int *dshell = mtile + c->nmesh + 1;
  // every tile writes all of its partials, and both int arrays are written
  // here, so nothing left by an earlier call is read
  HIPCHK(c, hipMemcpyAsync(mtile, c->diag_mtile.data(), (size_t)(c->nmesh + 1) * sizeof(int),
                           hipMemcpyHostToDevice, s));
  HIPCHK(c, hipMemcpyAsync(dshell, shell, (size_t)c->nmesh * sizeof(int), hipMemcpyHostToDevice, s));
  LCHK(c, launch_diagnostics(c->L, c->conts[md].kf(), dshell, part, mtile, c->nmesh, out, s));
  HIPCHK(c, hipStreamSynchronize(s));
  return PANGU_OK;
}
