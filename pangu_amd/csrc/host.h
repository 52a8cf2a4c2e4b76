// Host-only helpers of pangu_hip (no HIP headers: these files are built
// with g++ -ffp-contract=off so host arithmetic is reproducible).
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/pangu_hip.h"

namespace pangu {
// Host block geometry (Parthenon UniformCartesian).
struct HostBlock {
  int nx[3], ng[3], n[3], s[3], e[3], lx[3], level;
  double xg[3], dx[3];
  int64_t cells;
  double xc(int d, int idx) const { return xg[d] + (idx + 0.5) * dx[d]; }
  double xf(int d, int idx) const { return xg[d] + idx * dx[d]; }
  int64_t id(int k, int j, int i) const { return ((int64_t)k * n[1] + j) * n[0] + i; }
};
HostBlock make_host_block(const pangu_block &b);

// Static mesh refinement (refine_host.cpp).  refined_mesh: any block of the
// context above level 0.  check_tiling: the blocks of each mesh tile its
// domain (no overlap, no gap); "" or the reason they do not.
// build_refinement_tables: gather tables of the cross-level ghost exchange
// and of flux correction; "" or why the mesh cannot be handled (neighbours
// more than one level apart, a prolongation stencil across a physical
// boundary or beyond the ghost zones).
struct XFill;
struct XProl;
struct XFlux;
bool refined_mesh(const std::vector<HostBlock> &hb);
std::string check_tiling(const std::vector<HostBlock> &hb, const std::vector<int> &bmesh,
                         const std::vector<pangu_config> &mcfg);
std::string build_refinement_tables(const std::vector<HostBlock> &hb, const std::vector<int> &bmesh,
                                    const std::vector<pangu_config> &mcfg, std::vector<XFill> &fill,
                                    std::vector<XProl> &prol, std::vector<XFlux> &flux);

namespace geom {
typedef double M4[4][4];
void code_metric(int kind, const double x[4], M4 out, double h, double a);
double det4(const M4 m);
bool inv4(const M4 m, M4 out);
void connection(int kind, const double x[4], const M4 gcon, double h, double a, double conn[64]);
// Fill one centre / face table entry; `out` points at component 0 of the
// entry and consecutive components are `stride` doubles apart.
void fill_centre(const pangu_config &cfg, const double x[4], double *out, int64_t stride);
void fill_face(const pangu_config &cfg, int d, const double x[4], double *out, int64_t stride);
// gcov[16], gcon[16], sqrt|g|, 1/sqrt(-g^00), sqrt(-1/g^00) at x (no connection).
void fill_metric(const pangu_config &cfg, const double x[4], double out[35]);
// fill_metric at n points (x1, x2) = x12[2i], x12[2i+1] (x0 = x3 = 0), OpenMP.
void fill_metric_points(const pangu_config &cfg, const double *x12, int n, double *out);
void fill_centre_table(const pangu_config &cfg, const HostBlock &H, int ni, int nj, int nk,
                       bool has3d, double *out);
void fill_face_table(const pangu_config &cfg, const HostBlock &H, int d, int fi, int fj, int fk,
                     bool has3d, double *out);
}  // namespace geom



// Initial data written into host arrays laid out [block][comp][cell]
// (entire ghosted blocks), for the "base" container.
struct ProblemParams {
  int problem;  // 1 Orszag-Tang (unit box), 2 FM torus (MKS), 3 FM torus (CKS)
  double rin, rmax, kappa, perturbation, beta, aphi_cut, fel_0;
  uint64_t seed;
  int mesh_nx[3];
  int mad;      // CKS torus field: 0 sane, 1 mad
};
int generate_problem(const pangu_config &cfg, const ProblemParams &pp,
                     const std::vector<HostBlock> &blocks, int np, int iBX, int iENT, int iKEL,
                     std::vector<double> &prim);

}  // namespace pangu
