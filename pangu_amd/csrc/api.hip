// C ABI of pangu_hip: context, geometry tables, container registration and
// the task entry points (see include/pangu_hip.h).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/pangu_hip.h"
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
  pangu_config cfg;
  int device = 0;
  int nblocks = 0, ndim = 1, np = 0, nc = 0, iBX = -1, iENT = 0, iKEL = -1;
  std::vector<HostBlock> hb;
  DevBlock *dblk = nullptr;
  DevGeo *dgeo = nullptr;
  std::vector<double *> geo_bufs;
  int *dnbr = nullptr;
  std::vector<DevContainer> conts;
  // shared tables
  double **dalfven = nullptr, **demf = nullptr, **dflag = nullptr, **dflag_rec = nullptr;
  uint8_t **dmask = nullptr;
  int32_t **dstatus = nullptr, **diters = nullptr;
  bool shared_ok = false;
  // scratch
  double *dred = nullptr;
  int *dcluster = nullptr;
  double *dest = nullptr;
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
        msg = "all meshblocks of a context must share one shape";
        return 0;
      }
      if (blocks[b].nx[d] > 1 && blocks[b].ng[d] < (cfg->limiter == PANGU_LIMITER_PPM4 ? 3 : 2)) {
        msg = "nghost too small for the limiter (ppm4 reads i-3..i+2)";
        return 0;
      }
    }
  return 1;
}

// Geometry sets: blocks with identical x1/x2 footprints (and, for x3-dependent
// metrics, identical x3 footprints) share one table set.
static int build_geometry(pangu_ctx *c, const pangu_block *blocks) {
  const pangu_config &cfg = c->cfg;
  const bool has3d = (cfg.metric == PANGU_METRIC_CKS) && c->ndim == 3;
  std::map<std::vector<double>, int> key2geo;
  std::vector<DevGeo> geos;
  std::vector<DevBlock> db(c->nblocks);
  for (int b = 0; b < c->nblocks; ++b) {
    const HostBlock &H = c->hb[b];
    DevBlock &D = db[b];
    for (int d = 0; d < 3; ++d) {
      D.n[d] = H.n[d];
      D.s[d] = H.s[d];
      D.e[d] = H.e[d];
      D.lx[d] = H.lx[d];
      D.dx[d] = H.dx[d];
    }
    D.ncell = H.cells;
    std::vector<double> key = {H.xg[0], H.dx[0], (double)H.n[0], H.xg[1], H.dx[1], (double)H.n[1]};
    if (has3d) {
      key.push_back(H.xg[2]);
      key.push_back(H.dx[2]);
      key.push_back((double)H.n[2]);
    } else {
      key.push_back(H.xc(2, H.s[2]));
    }
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
    std::vector<double> cen((size_t)(G.ncen * CG_NCOMP));
    geom::fill_centre_table(cfg, H, G.ni, G.nj, G.nk, has3d, cen.data());
    double *dcen = nullptr;
    if (upload(c, &dcen, cen.data(), cen.size())) return PANGU_EDEVICE;
    c->geo_bufs.push_back(dcen);
    G.cen = dcen;
    const int nfd = has3d ? 3 : 2;
    for (int d = 0; d < 3; ++d) {
      G.fi[d] = G.ni + (d == 0 ? 1 : 0);
      G.fj[d] = G.nj + (d == 1 ? 1 : 0);
      if (d >= c->ndim || d >= nfd) {
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
  c->L.hblk = db[0];
  return 0;
}

// Neighbour table for same-level exchange: 27 entries per block
// (offset (o0,o1,o2) in {-1,0,1}^3), -1 where a physical boundary is.
static int build_neighbours(pangu_ctx *c) {
  const int *nb = c->cfg.nb;
  std::vector<int> nbr((size_t)c->nblocks * 27, -1);
  std::map<std::vector<int>, int> at;
  for (int b = 0; b < c->nblocks; ++b) at[{c->hb[b].lx[0], c->hb[b].lx[1], c->hb[b].lx[2]}] = b;
  for (int b = 0; b < c->nblocks; ++b)
    for (int o2 = -1; o2 <= 1; ++o2)
      for (int o1 = -1; o1 <= 1; ++o1)
        for (int o0 = -1; o0 <= 1; ++o0) {
          const int o[3] = {o0, o1, o2};
          std::vector<int> lx(3);
          bool ok = true;
          for (int d = 0; d < 3; ++d) {
            lx[d] = c->hb[b].lx[d] + o[d];
            if (lx[d] < 0 || lx[d] >= nb[d]) {
              if (c->cfg.bc[d][0] == PANGU_BC_PERIODIC)
                lx[d] = (lx[d] + nb[d]) % nb[d];
              else
                ok = false;
            }
          }
          if (!ok) continue;
          auto it = at.find(lx);
          if (it != at.end()) nbr[(size_t)b * 27 + (o2 + 1) * 9 + (o1 + 1) * 3 + (o0 + 1)] = it->second;
        }
  return upload(c, &c->dnbr, nbr.data(), nbr.size());
}

namespace pangu {
void set_kernel_clock(pangu_ctx *c, KernelClock *clock) {
  if (c) c->L.clock = clock;
}
}  // namespace pangu

extern "C" {

int pangu_create(const pangu_config *cfg, const pangu_block *blocks, int nblocks, int device,
                 pangu_ctx **out) {
  if (!out) return PANGU_EINVAL;
  *out = nullptr;
  std::string msg;
  if (!validate(cfg, blocks, nblocks, msg)) {
    fprintf(stderr, "pangu_create: %s\n", msg.c_str());
    return PANGU_EINVAL;
  }
  pangu_ctx *c = new pangu_ctx();
  c->cfg = *cfg;
  c->device = device;
  c->nblocks = nblocks;
  if (hipSetDevice(device) != hipSuccess) {
    delete c;
    return PANGU_EDEVICE;
  }
  c->ndim = (blocks[0].nx[2] > 1) ? 3 : (blocks[0].nx[1] > 1 ? 2 : 1);
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
  const int nplanes = c->hb[0].nx[2];
  if (!r) r = upload(c, &c->dred, std::vector<double>((size_t)nblocks * nplanes * 3, 0.0).data(),
                     (size_t)nblocks * nplanes * 3);
  if (!r) r = upload(c, &c->dcluster, std::vector<int>((size_t)nblocks * nplanes, 0).data(),
                     (size_t)nblocks * nplanes);
  if (!r) r = upload(c, &c->dest, std::vector<double>(1, 0.0).data(), 1);
  if (r) {
    fprintf(stderr, "pangu_create: %s\n", c->err.c_str());
    pangu_destroy(c);
    return r;
  }
  KCtx &k = c->L.k;
  k.blk = c->dblk;
  k.geo = c->dgeo;
  k.nblocks = nblocks;
  k.np = c->np;
  k.nc = c->nc;
  k.iBX = c->iBX;
  k.iENT = c->iENT;
  k.iKEL = c->iKEL;
  k.ndim = c->ndim;
  k.nk = 1;
  k.nkg = 1;
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
  for (int d = 0; d < 3; ++d) {
    c->L.nb[d] = cfg->nb[d];
    c->L.bc[d][0] = cfg->bc[d][0];
    c->L.bc[d][1] = cfg->bc[d][1];
  }
  *out = c;
  return PANGU_OK;
}

int pangu_destroy(pangu_ctx *c) {
  if (!c) return PANGU_OK;
  hipSetDevice(c->device);
  for (double *p : c->geo_bufs) hipFree(p);
  for (auto &ct : c->conts) {
    hipFree(ct.dprim);
    hipFree(ct.dcons);
    for (int d = 0; d < 3; ++d) hipFree(ct.dflux[d]);
  }
  void *ptrs[] = {c->dblk, c->dgeo, c->dnbr, c->dalfven, c->demf, c->dflag, c->dflag_rec,
                  c->dmask, c->dstatus, c->diters, c->dred, c->dcluster, c->dest};
  for (void *p : ptrs)
    if (p) hipFree(p);
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

int pangu_set_split_source(pangu_ctx *c, int split) {
  if (!c) return PANGU_EINVAL;
  c->L.split_source = split != 0;
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
  hipSetDevice(c->device);
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
  hipSetDevice(c->device);
  const size_t nb = (size_t)c->nblocks;
  if (upload(c, &c->dalfven, s->alfven, nb * 3)) return PANGU_EDEVICE;
  if (s->emf && upload(c, &c->demf, s->emf, nb * 3)) return PANGU_EDEVICE;
  if (s->flag && upload(c, &c->dflag, s->flag, nb)) return PANGU_EDEVICE;
  if (s->flag_recovery && upload(c, &c->dflag_rec, s->flag_recovery, nb)) return PANGU_EDEVICE;
  if (s->mask && upload(c, &c->dmask, s->mask, nb)) return PANGU_EDEVICE;
  if (s->c2p_status && upload(c, &c->dstatus, s->c2p_status, nb)) return PANGU_EDEVICE;
  if (s->c2p_iters && upload(c, &c->diters, s->c2p_iters, nb)) return PANGU_EDEVICE;
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

int pangu_calculate_fluxes(pangu_ctx *c, int md, void *stream) {
  if (int r = check_md(c, md)) return r;
  for (int d = 0; d < c->ndim; ++d)
    if (!c->conts[md].dflux[d]) return fail(c, PANGU_EINVAL, "container has no flux tables");
  LCHK(c, launch_fluxes(c->L, c->conts[md].kf(), c->dalfven, STREAM(stream)));
  return PANGU_OK;
}

int pangu_constrained_transport(pangu_ctx *c, int md, void *stream) {
  if (int r = check_md(c, md)) return r;
  if (!c->cfg.enable_B) return PANGU_OK;
  LCHK(c, launch_ct(c->L, c->conts[md].kf(), c->demf, STREAM(stream)));
  return PANGU_OK;
}

int pangu_flux_divergence(pangu_ctx *c, int md_in, int md_dudt, void *stream) {
  if (int r = check_md(c, md_in)) return r;
  if (int r = check_md(c, md_dudt)) return r;
  const KFields f = c->conts[md_in].kf();
  LCHK(c, launch_div_update(c->L, 1, f, f, f, c->conts[md_dudt].dcons, nullptr, 0.0, 0.0,
                            STREAM(stream)));
  return PANGU_OK;
}

int pangu_update_independent(pangu_ctx *c, int md_init, int md_dudt, double beta_dt, int md_out,
                             void *stream) {
  if (int r = check_md(c, md_init)) return r;
  if (int r = check_md(c, md_dudt)) return r;
  if (int r = check_md(c, md_out)) return r;
  LCHK(c, launch_div_update(c->L, 2, c->conts[md_init].kf(), c->conts[md_init].kf(),
                            c->conts[md_out].kf(), c->conts[md_dudt].dcons, nullptr, 0.0, beta_dt,
                            STREAM(stream)));
  return PANGU_OK;
}

int pangu_add_geometric_source(pangu_ctx *c, int md, double beta_dt, void *stream) {
  if (int r = check_md(c, md)) return r;
  const KFields f = c->conts[md].kf();
  LCHK(c, launch_div_update(c->L, 3, f, f, f, nullptr, nullptr, 0.0, beta_dt, STREAM(stream)));
  return PANGU_OK;
}

static SharedTabs shared_tabs(pangu_ctx *c) {
  SharedTabs S;
  S.flag = c->dflag;
  S.flag_rec = c->dflag_rec;
  S.status = c->dstatus;
  S.iters = c->diters;
  return S;
}

int pangu_recovery(pangu_ctx *c, int md, void *stream) {
  if (int r = check_md(c, md)) return r;
  LCHK(c, launch_recover(c->L, c->conts[md].kf(), shared_tabs(c), c->cfg.recovery, false,
                         STREAM(stream)));
  return PANGU_OK;
}

int pangu_apply_electron_heating(pangu_ctx *c, int md, void *stream) {
  if (int r = check_md(c, md)) return r;
  if (!c->cfg.enable_heating) return PANGU_OK;
  LCHK(c, launch_recover(c->L, c->conts[md].kf(), shared_tabs(c), -1, true, STREAM(stream)));
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
  HIPCHK(c, hipMemcpyAsync(dt_out, c->dest, sizeof(double), hipMemcpyDeviceToHost, STREAM(stream)));
  HIPCHK(c, hipStreamSynchronize(STREAM(stream)));
  return PANGU_OK;
}

int pangu_exchange_ghosts(pangu_ctx *c, int md, void *stream) {
  if (int r = check_md(c, md)) return r;
  LCHK(c, launch_exchange(c->L, c->conts[md].kf(), c->dflag, c->dnbr, STREAM(stream)));
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
  if (!dt_state) return fail(c, PANGU_EINVAL, "dt_state is null");
  hipStream_t s = STREAM(stream);
  const KFields f0 = c->conts[md0].kf(), fi = c->conts[md_init].kf(), f1 = c->conts[md1].kf();
  LCHK(c, launch_fix_cons(c->L, f0, c->dmask, true, true, s));
  LCHK(c, launch_fluxes(c->L, f0, c->dalfven, s));
  if (c->cfg.enable_B) LCHK(c, launch_ct(c->L, f0, c->demf, s));
  if (c->L.split_source && c->cfg.metric != PANGU_METRIC_MINKOWSKI) {
    LCHK(c, launch_div_update(c->L, 5, f0, fi, f1, nullptr, dt_state, beta, 0.0, s));
    LCHK(c, launch_div_update(c->L, 3, f0, fi, f1, nullptr, dt_state, beta, 0.0, s));
  } else {
    LCHK(c, launch_div_update(c->L, 4, f0, fi, f1, nullptr, dt_state, beta, 0.0, s));
  }
  if (exchange_after == 0) LCHK(c, launch_exchange(c->L, f1, c->dflag, c->dnbr, s));
  LCHK(c, launch_recover(c->L, f1, shared_tabs(c), c->cfg.recovery, c->cfg.enable_heating != 0, s));
  if (c->cfg.recovery != PANGU_RECOVERY_ROBUSTER)
    LCHK(c, launch_fix_recovery(c->L, f1, c->dflag, c->dcluster, s));
  if (exchange_after != 0) LCHK(c, launch_exchange(c->L, f1, c->dflag, c->dnbr, s));
  LCHK(c, launch_bc(c->L, f1, c->dflag, s));
  if (last_stage) LCHK(c, launch_dt(c->L, c->dalfven, c->dred, dt_state, dt_factor, nullptr, s));
  return PANGU_OK;
}

}  // extern "C"
