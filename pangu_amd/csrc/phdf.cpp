// Parthenon-style HDF5 output of cell fields (include/pangu_phdf.h).  Host
// code over the HDF5 1.10 C API; built into its own library so the GPU
// library carries no HDF5 dependency.
#include <hdf5.h>

#include <cstdio>
#include <string>
#include <vector>

#include "../../include/pangu_phdf.h"

namespace {

std::string g_err;

int fail(const std::string &m) {
  g_err = m;
  return -1;
}

// RAII holder of an HDF5 identifier.
struct H {
  hid_t id;
  herr_t (*close)(hid_t);
  H(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
  ~H() {
    if (id >= 0) close(id);
  }
  operator hid_t() const { return id; }
};

bool attr(hid_t loc, const char *name, hid_t mem, hid_t file_type, const void *v, hsize_t n) {
  H sp(H5Screate_simple(1, &n, nullptr), H5Sclose);
  H a(H5Acreate2(loc, name, file_type, sp, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
  return sp >= 0 && a >= 0 && H5Awrite(a, mem, v) >= 0;
}
bool attr_i(hid_t loc, const char *name, const int *v, hsize_t n) {
  return attr(loc, name, H5T_NATIVE_INT, H5T_STD_I32LE, v, n);
}
bool attr_d(hid_t loc, const char *name, const double *v, hsize_t n) {
  return attr(loc, name, H5T_NATIVE_DOUBLE, H5T_IEEE_F64LE, v, n);
}
// Array of fixed-length strings (length of the longest + 1).
bool attr_s(hid_t loc, const char *name, const std::vector<std::string> &v) {
  size_t len = 1;
  for (auto &s : v) len = s.size() + 1 > len ? s.size() + 1 : len;
  std::vector<char> buf(len * v.size(), 0);
  for (size_t i = 0; i < v.size(); ++i) v[i].copy(&buf[i * len], v[i].size());
  H t(H5Tcopy(H5T_C_S1), H5Tclose);
  if (t < 0 || H5Tset_size(t, len) < 0) return false;
  return attr(loc, name, t, t, buf.data(), v.size());
}

bool dataset(hid_t loc, const char *name, hid_t mem, hid_t file_type, const void *v, int rank,
             const hsize_t *dims) {
  H sp(H5Screate_simple(rank, dims, nullptr), H5Sclose);
  H d(H5Dcreate2(loc, name, file_type, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
  return sp >= 0 && d >= 0 && H5Dwrite(d, mem, H5S_ALL, H5S_ALL, H5P_DEFAULT, v) >= 0;
}

std::string base_name(const std::string &p) {
  const size_t s = p.find_last_of('/');
  return s == std::string::npos ? p : p.substr(s + 1);
}

// XDMF 2 sidecar: one 3DRectMesh grid per block, cell attributes as
// hyperslabs (block b, component c) of the HDF5 datasets.
bool write_xdmf(const std::string &path, const pangu_phdf_mesh &m, int nvars,
                const pangu_phdf_var *vars) {
  FILE *f = std::fopen((path + ".xdmf").c_str(), "w");
  if (!f) return false;
  const std::string h5 = base_name(path);
  const int nb = m.nblocks, nx = m.nx[0], ny = m.nx[1], nz = m.nx[2];
  std::fprintf(f, "<?xml version=\"1.0\" ?>\n<!DOCTYPE Xdmf SYSTEM \"Xdmf.dtd\" []>\n");
  std::fprintf(f, "<Xdmf Version=\"2.0\">\n<Domain>\n");
  std::fprintf(f, "<Grid Name=\"Mesh\" GridType=\"Collection\">\n");
  std::fprintf(f, "  <Time Value=\"%.17g\"/>\n", m.time);
  const char *axes[3] = {"x", "y", "z"};
  for (int b = 0; b < nb; ++b) {
    std::fprintf(f, "  <Grid GridType=\"Uniform\" Name=\"block%d\">\n", b);
    std::fprintf(f, "    <Topology TopologyType=\"3DRectMesh\" Dimensions=\"%d %d %d\"/>\n", nz + 1,
                 ny + 1, nx + 1);
    std::fprintf(f, "    <Geometry GeometryType=\"VXVYVZ\">\n");
    for (int d = 0; d < 3; ++d) {
      const int n = m.nx[d] + 1;
      std::fprintf(f,
                   "      <DataItem ItemType=\"HyperSlab\" Dimensions=\"%d\">\n"
                   "        <DataItem Dimensions=\"3 2\" NumberType=\"Int\" Format=\"XML\">%d 0 1 1 1 %d</DataItem>\n"
                   "        <DataItem Dimensions=\"%d %d\" NumberType=\"Float\" Precision=\"8\" Format=\"HDF\">%s:/Locations/%s</DataItem>\n"
                   "      </DataItem>\n",
                   n, b, n, nb, n, h5.c_str(), axes[d]);
    }
    std::fprintf(f, "    </Geometry>\n");
    for (int v = 0; v < nvars; ++v) {
      const int nc = vars[v].ncomp;
      for (int c = 0; c < nc; ++c) {
        std::string nm = vars[v].name;
        if (nc > 1) nm += "_" + std::to_string(c);
        std::fprintf(f, "    <Attribute Name=\"%s\" Center=\"Cell\">\n", nm.c_str());
        if (nc == 1)
          std::fprintf(f,
                       "      <DataItem ItemType=\"HyperSlab\" Dimensions=\"%d %d %d\">\n"
                       "        <DataItem Dimensions=\"3 4\" NumberType=\"Int\" Format=\"XML\">%d 0 0 0 1 1 1 1 1 %d %d %d</DataItem>\n"
                       "        <DataItem Dimensions=\"%d %d %d %d\" NumberType=\"Float\" Precision=\"8\" Format=\"HDF\">%s:/%s</DataItem>\n",
                       nz, ny, nx, b, nz, ny, nx, nb, nz, ny, nx, h5.c_str(), vars[v].name);
        else
          std::fprintf(f,
                       "      <DataItem ItemType=\"HyperSlab\" Dimensions=\"%d %d %d\">\n"
                       "        <DataItem Dimensions=\"3 5\" NumberType=\"Int\" Format=\"XML\">%d %d 0 0 0 1 1 1 1 1 1 1 %d %d %d</DataItem>\n"
                       "        <DataItem Dimensions=\"%d %d %d %d %d\" NumberType=\"Float\" Precision=\"8\" Format=\"HDF\">%s:/%s</DataItem>\n",
                       nz, ny, nx, b, c, nz, ny, nx, nb, nc, nz, ny, nx, h5.c_str(), vars[v].name);
        std::fprintf(f, "      </DataItem>\n    </Attribute>\n");
      }
    }
    std::fprintf(f, "  </Grid>\n");
  }
  std::fprintf(f, "</Grid>\n</Domain>\n</Xdmf>\n");
  return std::fclose(f) == 0;
}

}  // namespace

extern "C" const char *pangu_phdf_last_error(void) { return g_err.c_str(); }

extern "C" int pangu_write_phdf(const char *path, const pangu_phdf_mesh *m, int nvars,
                                const pangu_phdf_var *vars) {
  if (!path || !m || m->nblocks <= 0 || !m->lx || !m->xmin || !m->dx || (nvars > 0 && !vars))
    return fail("phdf: null argument or no blocks");
  for (int d = 0; d < 3; ++d)
    if (m->nx[d] < 1 || m->root_nx[d] < 1) return fail("phdf: block and mesh sizes must be >= 1");
  for (int v = 0; v < nvars; ++v)
    if (!vars[v].name || !vars[v].data || vars[v].ncomp < 1) return fail("phdf: bad variable");
  H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
  H file(H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
  if (file < 0) return fail(std::string("phdf: cannot create ") + path);
  const int nb = m->nblocks;
  const hsize_t hnb = (hsize_t)nb;

  {  // Info
    H info(H5Gcreate2(file, "Info", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
    const int one_i[1] = {0};
    const int ver = 3;
    int max_level = 0;
    for (int b = 0; m->level && b < nb; ++b)
      if (m->level[b] > max_level) max_level = m->level[b];
    bool ok = info >= 0 && attr_i(info, "OutputFormatVersion", &ver, 1) &&
              attr_i(info, "NCycle", &m->ncycle, 1) && attr_d(info, "Time", &m->time, 1) &&
              attr_d(info, "dt", &m->dt, 1) && attr_i(info, "NumDims", &m->ndim, 1) &&
              attr_i(info, "NumMeshBlocks", &nb, 1) && attr_i(info, "MeshBlockSize", m->nx, 3) &&
              attr_i(info, "RootGridSize", m->root_nx, 3) && attr_i(info, "IncludesGhost", one_i, 1) &&
              attr_i(info, "NGhost", &m->ng, 1) && attr_i(info, "MaxLevel", &max_level, 1) &&
              attr_s(info, "Coordinates", {"UniformCartesian"});
    double dom[9];
    for (int d = 0; d < 3; ++d) {
      dom[3 * d] = m->root_min[d];
      dom[3 * d + 1] = m->root_max[d];
      dom[3 * d + 2] = 1.0;
    }
    ok = ok && attr_d(info, "RootGridDomain", dom, 9);
    std::vector<std::string> names;
    std::vector<int> ncomp;
    for (int v = 0; v < nvars; ++v) {
      names.push_back(vars[v].name);
      ncomp.push_back(vars[v].ncomp);
    }
    if (nvars > 0)
      ok = ok && attr_s(info, "OutputDatasetNames", names) &&
           attr_i(info, "NumComponents", ncomp.data(), (hsize_t)nvars);
    if (!ok) return fail("phdf: cannot write /Info");
  }
  {  // block locations
    const hsize_t d2[2] = {hnb, 3};
    H blocks(H5Gcreate2(file, "Blocks", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
    std::vector<int> levels(nb, 0);
    if (m->level) levels.assign(m->level, m->level + nb);
    if (blocks < 0 || !dataset(blocks, "loc.lx123", H5T_NATIVE_INT64, H5T_STD_I64LE, m->lx, 2, d2) ||
        !dataset(file, "LogicalLocations", H5T_NATIVE_INT64, H5T_STD_I64LE, m->lx, 2, d2) ||
        !dataset(file, "Levels", H5T_NATIVE_INT, H5T_STD_I32LE, levels.data(), 1, &hnb))
      return fail("phdf: cannot write block locations");
    H vl(H5Gcreate2(file, "VolumeLocations", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
    H fl(H5Gcreate2(file, "Locations", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
    if (vl < 0 || fl < 0) return fail("phdf: cannot create location groups");
    const char *axes[3] = {"x", "y", "z"};
    for (int d = 0; d < 3; ++d) {
      const int n = m->nx[d];
      std::vector<double> cen((size_t)nb * n), face((size_t)nb * (n + 1));
      for (int b = 0; b < nb; ++b) {
        const double x0 = m->xmin[3 * b + d], h = m->dx[3 * b + d];
        for (int i = 0; i < n; ++i) cen[(size_t)b * n + i] = x0 + (i + 0.5) * h;
        for (int i = 0; i <= n; ++i) face[(size_t)b * (n + 1) + i] = x0 + i * h;
      }
      const hsize_t dc[2] = {hnb, (hsize_t)n}, df[2] = {hnb, (hsize_t)n + 1};
      if (!dataset(vl, axes[d], H5T_NATIVE_DOUBLE, H5T_IEEE_F64LE, cen.data(), 2, dc) ||
          !dataset(fl, axes[d], H5T_NATIVE_DOUBLE, H5T_IEEE_F64LE, face.data(), 2, df))
        return fail("phdf: cannot write cell locations");
    }
  }
  for (int v = 0; v < nvars; ++v) {
    const hsize_t nz = (hsize_t)m->nx[2], ny = (hsize_t)m->nx[1], nx = (hsize_t)m->nx[0];
    bool ok;
    if (vars[v].ncomp == 1) {
      const hsize_t d4[4] = {hnb, nz, ny, nx};
      ok = dataset(file, vars[v].name, H5T_NATIVE_DOUBLE, H5T_IEEE_F64LE, vars[v].data, 4, d4);
    } else {
      const hsize_t d5[5] = {hnb, (hsize_t)vars[v].ncomp, nz, ny, nx};
      ok = dataset(file, vars[v].name, H5T_NATIVE_DOUBLE, H5T_IEEE_F64LE, vars[v].data, 5, d5);
    }
    if (!ok) return fail(std::string("phdf: cannot write variable ") + vars[v].name);
  }
  if (!write_xdmf(path, *m, nvars, vars)) return fail("phdf: cannot write the XDMF sidecar");
  return 0;
}
