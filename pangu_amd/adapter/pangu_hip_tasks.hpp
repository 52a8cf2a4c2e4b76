// Parthenon-side adapter of the pangu_hip C ABI: the task functions PANGU's
// task list adds (pangu/src/task_list/ideal_grmhd.cc:62-110), with the
// reference's signatures, running on the device buffers of each MeshData.
//
// Use: call pangu_hip_tasks::Setup(pin, base_md, device) once per partition
// after the mesh is built (and again after remeshing), then add these
// functions with tl.AddTask where ideal_grmhd.cc adds the originals.
// Parthenon's own FluxDivergence, UpdateIndependentData, boundary exchange
// and ApplyBoundaryConditionsMD stay in place: they act on the same arrays,
// and Kokkos orders them on the same stream.
#pragma once

#include <map>
#include <string>
#include <vector>

#include <parthenon/package.hpp>

#include "pangu_hip.h"

namespace pangu_hip_tasks {

using parthenon::MeshData;
using parthenon::ParameterInput;
using parthenon::Real;
using parthenon::TaskStatus;

// One context per partition; containers registered by stage name.
struct Partition {
  pangu_ctx *ctx = nullptr;
  std::map<std::string, int> handle;  // "base", "1", "dUdt" -> registration handle
};

inline std::map<int, Partition> &Registry() {
  static std::map<int, Partition> r;
  return r;
}

// Index of s in the list, or a failure naming the parameter, as the
// reference rejects unknown choices (package_registration.cc:38-40,
// riemann_solver.cc:25, heating_model.h:118).
inline int Choice(const std::string &what, const std::string &s, const std::vector<std::string> &opts) {
  for (size_t i = 0; i < opts.size(); ++i)
    if (opts[i] == s) return (int)i;
  PARTHENON_FAIL(("unknown " + what + ": " + s).c_str());
  return -1;
}

// The <core>, <electron> and <metric> blocks with the reference's names and
// defaults (package_registration.cc:21-57,158-167).  recovery = "onedvsq"
// (1Dvsq alone) is this library's extension; the reference accepts
// "chainer" and "robuster".
inline pangu_config MakeConfig(ParameterInput *pin) {
  pangu_config c{};
  const std::string mt = pin->GetOrAddString("metric", "type", "minkowski");
  c.metric = Choice("metric type", mt, {"minkowski", "bl", "cks", "mks"});  // PANGU_METRIC_* order
  c.a = pin->GetOrAddReal("metric", "a", 0.0);
  c.h = pin->GetOrAddReal("metric", "h", 0.0);
  c.cfl = pin->GetOrAddReal("core", "cfl_number", 0.8);
  c.gamma = pin->GetOrAddReal("core", "adiabatic_index", 5. / 3.);
  c.riemann = Choice("riemann_solver", pin->GetOrAddString("core", "riemann_solver", "laxf"),
                     {"laxf", "hll"});
  c.limiter = Choice("limiter", pin->GetOrAddString("core", "limiter", "ppm4"), {"mc", "ppm4"});
  c.density_floor = pin->GetOrAddReal("core", "density_floor", 1.0e-6);
  c.energy_floor = pin->GetOrAddReal("core", "energy_floor", 1.0e-8);
  c.density_floor_pow = pin->GetOrAddReal("core", "density_floor_pow", -1.5);
  c.energy_floor_pow = pin->GetOrAddReal("core", "energy_floor_pow", -2.5);
  c.enable_B = pin->GetOrAddBoolean("core", "enable_B", false);
  c.recovery = Choice("recovery", pin->GetOrAddString("core", "recovery", "chainer"),
                      {"chainer", "robuster", "onedvsq"});
  c.enable_heating = pin->GetOrAddBoolean("core", "enable_heating", false);
  c.sigma_max = pin->GetOrAddReal("core", "sigma_max", 50);
  c.lorentz_max = pin->GetOrAddReal("core", "lorentz_max", 50);
  c.heat_model = Choice("electron model", pin->GetOrAddString("electron", "model", "constant"),
                        {"constant", "howes", "kawazura", "werner", "rowan", "sharma"});
  c.fel_constant = pin->GetOrAddReal("electron", "fel_constant", 0.1);
  c.gamma_e = pin->GetOrAddReal("electron", "gamma_e", 4. / 3.);
  c.gamma_p = pin->GetOrAddReal("electron", "gamma_p", 5. / 3.);
  c.limit_kel = pin->GetOrAddBoolean("electron", "limit_kel", true);
  c.suppress_highb_heat = pin->GetOrAddBoolean("electron", "suppress_highb_heat", false);
  c.enforce_positive_dissipation = pin->GetOrAddBoolean("electron", "enforce_positive_dissipation", false);
  c.ratio_min = pin->GetOrAddReal("electron", "ratio_min", 0.001);
  c.ratio_max = pin->GetOrAddReal("electron", "ratio_max", 1000.0);
  for (int d = 0; d < 3; ++d) {
    c.nb[d] = pin->GetOrAddInteger("parthenon/mesh", "nx" + std::to_string(d + 1), 1) /
              pin->GetOrAddInteger("parthenon/meshblock", "nx" + std::to_string(d + 1), 1);
    for (int s = 0; s < 2; ++s) {
      const std::string key = std::string(s ? "o" : "i") + "x" + std::to_string(d + 1) + "_bc";
      c.bc[d][s] = Choice(key, pin->GetOrAddString("parthenon/mesh", key, "periodic"),
                          {"periodic", "outflow", "reflecting"});
    }
  }
  return c;
}

// Block geometry of a MeshData (UniformCartesian coordinates, logical
// locations for refined meshes).
inline std::vector<pangu_block> MakeBlocks(MeshData<Real> *md) {
  std::vector<pangu_block> out;
  for (int b = 0; b < md->NumBlocks(); ++b) {
    auto pmb = md->GetBlockData(b)->GetBlockPointer();
    const auto &co = pmb->coords;
    pangu_block B{};
    for (int d = 0; d < 3; ++d) {
      B.nx[d] = pmb->block_size.nx(d);
      B.ng[d] = B.nx[d] > 1 ? parthenon::Globals::nghost : 0;
      B.lx[d] = pmb->loc.l(d);
      B.xmin[d] = pmb->block_size.xmin(d);
      B.dx[d] = co.Dxc(d + 1);
    }
    B.level = pmb->loc.level();
    out.push_back(B);
  }
  return out;
}

// The primitive and conservative packs of one container, in the
// registration order of package_registration.cc:84-96.
inline int RegisterContainer(pangu_ctx *ctx, MeshData<Real> *md) {
  auto prim = md->PackVariables(std::vector<std::string>{"primitive"});
  auto cons = md->PackVariablesAndFluxes(std::vector<std::string>{"conservative"});
  const int nb = prim.GetDim(5), np = prim.GetDim(4), nc = cons.GetDim(4);
  std::vector<double *> hp(nb * np), hc(nb * nc), hf[3];
  for (int b = 0; b < nb; ++b) {
    for (int l = 0; l < np; ++l) hp[b * np + l] = prim(b, l).data();
    for (int l = 0; l < nc; ++l) hc[b * nc + l] = cons(b, l).data();
  }
  for (int d = 0; d < 3; ++d) {
    hf[d].resize(nb * nc);
    for (int b = 0; b < nb; ++b)
      for (int l = 0; l < nc; ++l) hf[d][b * nc + l] = cons(b).flux(d + 1, l).data();
  }
  pangu_container c{hp.data(), hc.data(), {hf[0].data(), hf[1].data(), hf[2].data()}};
  int handle = -1;
  return pangu_register_container(ctx, &c, &handle) == PANGU_OK ? handle : -1;
}

// The OneCopy fields (package_registration.cc:117-144) of the base container.
inline int RegisterShared(pangu_ctx *ctx, MeshData<Real> *md) {
  auto one = [&](const char *name, std::vector<double *> &v) {
    auto p = md->PackVariables(std::vector<std::string>{name});
    const int nb = p.GetDim(5), nl = p.GetDim(4);
    v.resize(nb * nl);
    for (int b = 0; b < nb; ++b)
      for (int l = 0; l < nl; ++l) v[b * nl + l] = p(b, l).data();
  };
  std::vector<double *> alfven, emf, flag;
  one("alfven", alfven);
  one("flag", flag);
  pangu_shared s{};
  s.alfven = alfven.data();
  s.flag = flag.data();
  if (md->GetBlockData(0)->HasVariable("electric_field")) {
    one("electric_field", emf);
    s.emf = emf.data();
  }
  return pangu_register_shared(ctx, &s);
}

// Creates the partition's context from its base container.
inline Partition &Setup(ParameterInput *pin, MeshData<Real> *base, int device) {
  Partition &p = Registry()[base->partition];
  if (p.ctx) pangu_destroy(p.ctx);
  p = Partition{};
  const pangu_config cfg = MakeConfig(pin);
  const std::vector<pangu_block> blocks = MakeBlocks(base);
  if (pangu_create(&cfg, blocks.data(), (int)blocks.size(), device, &p.ctx) != PANGU_OK)
    PARTHENON_FAIL("pangu_create failed");
  if (RegisterShared(p.ctx, base) != PANGU_OK) PARTHENON_FAIL(pangu_last_error(p.ctx));
  return p;
}

inline Partition &Part(MeshData<Real> *md) { return Registry().at(md->partition); }

// The container's handle, registered on first use.
inline int Handle(MeshData<Real> *md) {
  Partition &p = Part(md);
  auto it = p.handle.find(md->StageName());
  if (it != p.handle.end()) return it->second;
  const int h = RegisterContainer(p.ctx, md);
  if (h < 0) PARTHENON_FAIL(pangu_last_error(p.ctx));
  p.handle[md->StageName()] = h;
  return h;
}

inline void *Stream() { return Kokkos::DefaultExecutionSpace().hip_stream(); }

inline TaskStatus Check(int r, pangu_ctx *c) {
  if (r != PANGU_OK) PARTHENON_FAIL(pangu_last_error(c));
  return TaskStatus::complete;
}

inline TaskStatus FixPrimitive(MeshData<Real> *md) {
  auto &p = Part(md);
  return Check(pangu_fix_primitive(p.ctx, Handle(md), Stream()), p.ctx);
}
inline TaskStatus CalculateConservative(MeshData<Real> *md) {
  auto &p = Part(md);
  return Check(pangu_calculate_conservative(p.ctx, Handle(md), Stream()), p.ctx);
}
inline TaskStatus CalculateFluxes(MeshData<Real> *md) {
  auto &p = Part(md);
  return Check(pangu_calculate_fluxes(p.ctx, Handle(md), Stream()), p.ctx);
}
inline TaskStatus ConstraintedTransport(MeshData<Real> *md) {
  auto &p = Part(md);
  return Check(pangu_constrained_transport(p.ctx, Handle(md), Stream()), p.ctx);
}
inline TaskStatus AddGeometricSource(MeshData<Real> *md, Real beta_dt) {
  auto &p = Part(md);
  return Check(pangu_add_geometric_source(p.ctx, Handle(md), beta_dt, Stream()), p.ctx);
}
inline TaskStatus Recovery(MeshData<Real> *md) {
  auto &p = Part(md);
  return Check(pangu_recovery(p.ctx, Handle(md), Stream()), p.ctx);
}
inline TaskStatus ApplyElectronHeating(MeshData<Real> *md) {
  auto &p = Part(md);
  return Check(pangu_apply_electron_heating(p.ctx, Handle(md), Stream()), p.ctx);
}
inline TaskStatus FixRecovery(MeshData<Real> *md) {
  auto &p = Part(md);
  return Check(pangu_fix_recovery(p.ctx, Handle(md), Stream()), p.ctx);
}
// EstimateTimestepMesh (initialization/timestep_estimation.cc): the new dt
// estimate of this partition (host value; waits for the stream).
inline Real EstimateTimestepMesh(MeshData<Real> *md) {
  auto &p = Part(md);
  double dt = 0.0;
  Check(pangu_estimate_timestep(p.ctx, Handle(md), &dt, Stream()), p.ctx);
  return dt;
}

}  // namespace pangu_hip_tasks
