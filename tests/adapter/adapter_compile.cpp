// Compiles the pangu_hip adapter against the interface stand-in and the C
// ABI, takes the address of every task function with the reference's
// signature (so each is instantiated and linked against libpangu_hip.so),
// and maps one input file's <core>/<electron>/<metric> blocks to a
// pangu_config.  Runs no device code.
#include <cstdio>

#include "pangu_hip_tasks.hpp"

using parthenon::MeshData;
using parthenon::Real;
using parthenon::TaskStatus;
using Task = TaskStatus (*)(MeshData<Real> *);

int main() {
  namespace T = pangu_hip_tasks;
  const Task tasks[] = {&T::FixPrimitive, &T::CalculateConservative, &T::CalculateFluxes,
                        &T::ConstraintedTransport, &T::Recovery, &T::ApplyElectronHeating,
                        &T::FixRecovery};
  TaskStatus (*source)(MeshData<Real> *, Real) = &T::AddGeometricSource;
  Real (*dt)(MeshData<Real> *) = &T::EstimateTimestepMesh;
  T::Partition &(*setup)(parthenon::ParameterInput *, MeshData<Real> *, int) = &T::Setup;
  parthenon::ParameterInput pin;
  pin.GetOrAddString("metric", "type", "mks");
  pin.GetOrAddReal("metric", "a", 0.9375);
  pin.GetOrAddString("core", "riemann_solver", "hll");
  pin.GetOrAddString("core", "recovery", "robuster");
  pin.GetOrAddString("electron", "model", "kawazura");
  pin.GetOrAddString("parthenon/mesh", "ix1_bc", "outflow");
  pin.GetOrAddInteger("parthenon/mesh", "nx1", 256);
  pin.GetOrAddInteger("parthenon/meshblock", "nx1", 64);
  const pangu_config c = T::MakeConfig(&pin);
  const bool ok = c.metric == PANGU_METRIC_MKS && c.a == 0.9375 && c.riemann == PANGU_RIEMANN_HLL &&
                  c.recovery == PANGU_RECOVERY_ROBUSTER && c.heat_model == PANGU_HEAT_KAWAZURA &&
                  c.bc[0][0] == PANGU_BC_OUTFLOW && c.bc[0][1] == PANGU_BC_PERIODIC && c.nb[0] == 4 &&
                  c.gamma == 5. / 3. && c.cfl == 0.8 && c.limiter == PANGU_LIMITER_PPM4;
  // an unknown choice fails as the reference's std::invalid_argument does
  bool rejected = false;
  try {
    parthenon::ParameterInput bad;
    bad.GetOrAddString("core", "recovery", "newton");
    (void)T::MakeConfig(&bad);
  } catch (const std::exception &) {
    rejected = true;
  }
  std::printf("tasks %zu source %d dt %d setup %d config %s unknown-choice %s\n",
              sizeof(tasks) / sizeof(tasks[0]), source != nullptr, dt != nullptr, setup != nullptr,
              ok ? "ok" : "wrong", rejected ? "rejected" : "accepted");
  return ok && rejected ? 0 : 1;
}
