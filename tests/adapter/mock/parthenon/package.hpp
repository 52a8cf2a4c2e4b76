// Compile-only stand-in for the parts of Parthenon's interface that the
// pangu_hip adapter (pangu_amd/adapter/pangu_hip_tasks.hpp) touches, with
// the member names the reference's sources use (MeshData::GetBlockData,
// MeshBlockData::GetBlockPointer, PackVariables, PackVariablesAndFluxes,
// GetDim, pack(b).flux(dir, l), ParameterInput::GetOrAdd*, TaskStatus).
// It lets a CPU test compile and link the adapter against libpangu_hip.so;
// it is not Parthenon and runs nothing.
#pragma once
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#define PARTHENON_FAIL(msg) throw std::runtime_error(msg)

namespace Kokkos {
struct DefaultExecutionSpace {
  void *hip_stream() const { return nullptr; }
};
}  // namespace Kokkos

namespace parthenon {
using Real = double;
enum class TaskStatus { fail, complete, incomplete, iterate };
namespace Globals {
inline int nghost = 3;
}

class ParameterInput {
 public:
  Real GetOrAddReal(const std::string &b, const std::string &n, Real v) { return get(b, n, v, reals_); }
  int GetOrAddInteger(const std::string &b, const std::string &n, int v) { return get(b, n, v, ints_); }
  bool GetOrAddBoolean(const std::string &b, const std::string &n, bool v) { return get(b, n, v, bools_); }
  std::string GetOrAddString(const std::string &b, const std::string &n, const std::string &v) {
    return get(b, n, v, strings_);
  }

 private:
  template <typename T>
  T get(const std::string &b, const std::string &n, T v, std::map<std::string, T> &m) {
    return m.emplace(b + "/" + n, v).first->second;
  }
  std::map<std::string, Real> reals_;
  std::map<std::string, int> ints_;
  std::map<std::string, bool> bools_;
  std::map<std::string, std::string> strings_;
};

struct ParArray3D {
  Real *p = nullptr;
  Real *data() const { return p; }
};

struct UniformCartesian {
  Real dx[3] = {1, 1, 1};
  Real Dxc(int dir) const { return dx[dir - 1]; }
};
struct RegionSize {
  int n[3] = {1, 1, 1};
  Real x0[3] = {0, 0, 0};
  int nx(int d) const { return n[d]; }
  Real xmin(int d) const { return x0[d]; }
};
struct LogicalLocation {
  int lx[3] = {0, 0, 0};
  int lev = 0;
  int l(int d) const { return lx[d]; }
  int level() const { return lev; }
};
struct MeshBlock {
  UniformCartesian coords;
  RegionSize block_size;
  LogicalLocation loc;
};

template <typename T>
struct MeshBlockData {
  std::shared_ptr<MeshBlock> block = std::make_shared<MeshBlock>();
  std::vector<std::string> vars;
  std::shared_ptr<MeshBlock> GetBlockPointer() const { return block; }
  bool HasVariable(const std::string &name) const {
    for (const auto &v : vars)
      if (v == name) return true;
    return false;
  }
};

struct BlockFluxes {
  std::vector<ParArray3D> f[3];
  ParArray3D flux(int dir, int l) const { return f[dir - 1][l]; }
};

struct VariablePack {
  int dims[6] = {1, 1, 1, 1, 1, 1};
  std::vector<std::vector<ParArray3D>> v;  // [block][component]
  int GetDim(int i) const { return dims[i - 1]; }
  ParArray3D operator()(int b, int l) const { return v[b][l]; }
};

struct VariableFluxPack : VariablePack {
  std::vector<BlockFluxes> fl;
  const BlockFluxes &operator()(int b) const { return fl[b]; }
  using VariablePack::operator();
};

template <typename T>
struct MeshData {
  int partition = 0;
  std::string stage = "base";
  std::vector<std::shared_ptr<MeshBlockData<T>>> blocks;
  int NumBlocks() const { return (int)blocks.size(); }
  std::shared_ptr<MeshBlockData<T>> GetBlockData(int b) const { return blocks[b]; }
  const std::string &StageName() const { return stage; }
  VariablePack PackVariables(const std::vector<std::string> &) { return VariablePack{}; }
  VariableFluxPack PackVariablesAndFluxes(const std::vector<std::string> &) { return VariableFluxPack{}; }
};
}  // namespace parthenon
