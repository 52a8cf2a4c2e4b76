// Block tree of a statically refined mesh, resolved on the host into the
// gather tables the refinement kernels run (refine.hip).
//
// A block is identified as in Parthenon's LogicalLocation by (level, lx):
// at level L a mesh has nb[d] << L blocks along each active dim, and the
// block at lx covers the level-L cells lx*nx .. lx*nx + nx - 1.  Active dims
// are a prefix (x1; x1, x2; x1, x2, x3).  Built with g++ like the other host
// sources; no floating point is involved.
#include <array>
#include <map>
#include <tuple>

#include "common.h"
#include "host.h"

namespace pangu {
namespace {

using Key = std::tuple<int, int, int64_t, int64_t, int64_t>;  // mesh, level, lx1, lx2, lx3

int64_t floor2(int64_t g) { return (g - (g & 1)) / 2; }

struct MeshInfo {
  int dim = 1;
  int nx[3] = {1, 1, 1};
  int nb[3] = {1, 1, 1};
  bool periodic[3] = {false, false, false};
};

struct Tree {
  const std::vector<HostBlock> &hb;
  const std::vector<int> &bmesh;
  std::vector<MeshInfo> mi;
  std::map<Key, int> at;
  std::string err;

  Tree(const std::vector<HostBlock> &h, const std::vector<int> &bm, const std::vector<pangu_config> &mcfg)
      : hb(h), bmesh(bm), mi(mcfg.size()) {
    std::vector<bool> seen(mcfg.size(), false);
    for (size_t b = 0; b < hb.size(); ++b) {
      const int m = bmesh[b];
      if (seen[m]) continue;
      seen[m] = true;
      MeshInfo &M = mi[m];
      M.dim = hb[b].n[2] > 1 ? 3 : (hb[b].n[1] > 1 ? 2 : 1);
      for (int d = 0; d < 3; ++d) {
        M.nx[d] = hb[b].nx[d];
        M.nb[d] = mcfg[m].nb[d];
        M.periodic[d] = mcfg[m].bc[d][0] == PANGU_BC_PERIODIC;
      }
    }
    for (size_t b = 0; b < hb.size(); ++b) {
      const Key k{bmesh[b], hb[b].level, hb[b].lx[0], hb[b].lx[1], hb[b].lx[2]};
      if (!at.emplace(k, (int)b).second && err.empty())
        err = "two blocks at one logical location (block " + std::to_string(b) + ")";
    }
  }
  // Cells of mesh m along d at level L.
  int64_t ncell(int m, int d, int L) const {
    return d < mi[m].dim ? ((int64_t)mi[m].nb[d] * mi[m].nx[d]) << L : 1;
  }
  // Wraps level-L cell index g along d into the domain; false outside a
  // non-periodic boundary.
  bool wrap(int m, int d, int L, int64_t &g) const {
    if (d >= mi[m].dim) {
      g = 0;
      return true;
    }
    const int64_t n = ncell(m, d, L);
    if (g >= 0 && g < n) return true;
    if (!mi[m].periodic[d]) return false;
    g = ((g % n) + n) % n;
    return true;
  }
  // Block at level L of mesh m that covers level-L cell g (wrapped), or -1.
  int find(int m, int L, const int64_t g[3]) const {
    if (L < 0) return -1;
    int64_t l[3];
    for (int d = 0; d < 3; ++d) l[d] = d < mi[m].dim ? g[d] / mi[m].nx[d] : 0;
    auto it = at.find(Key{m, L, l[0], l[1], l[2]});
    return it == at.end() ? -1 : it->second;
  }
  // The leaf covering level-L cell g: at level L, L - 1 or L + 1 (lev).
  int cover(int m, int L, const int64_t g[3], int &lev) const {
    int b = find(m, L, g);
    if (b >= 0) {
      lev = L;
      return b;
    }
    int64_t c[3], f[3];
    for (int d = 0; d < 3; ++d) {
      c[d] = d < mi[m].dim ? (g[d] >> 1) : 0;
      f[d] = d < mi[m].dim ? 2 * g[d] : 0;
    }
    b = find(m, L - 1, c);
    if (b >= 0) {
      lev = L - 1;
      return b;
    }
    b = find(m, L + 1, f);
    if (b >= 0) lev = L + 1;
    return b;
  }
  // Index in block B's ghosted array of its mesh's level cell g along d.
  int local(const HostBlock &B, int d, int64_t g) const {
    return B.n[d] > 1 ? (int)(g - (int64_t)B.lx[d] * B.nx[d] + B.s[d]) : B.s[d];
  }
  int cell(const HostBlock &B, const int64_t g[3]) const {
    return (int)B.id(local(B, 2, g[2]), local(B, 1, g[1]), local(B, 0, g[0]));
  }
};

}  // namespace

bool refined_mesh(const std::vector<HostBlock> &hb) {
  for (const HostBlock &B : hb)
    if (B.level != 0) return true;
  return false;
}

std::string check_tiling(const std::vector<HostBlock> &hb, const std::vector<int> &bmesh,
                         const std::vector<pangu_config> &mcfg) {
  Tree T(hb, bmesh, mcfg);
  if (!T.err.empty()) return T.err;
  const int nmesh = (int)mcfg.size();
  std::vector<int> lmax(nmesh, 0);
  for (size_t b = 0; b < hb.size(); ++b) {
    const HostBlock &B = hb[b];
    const int m = bmesh[b];
    if (B.level < 0 || B.level > 20) return "refinement level out of range (0..20)";
    for (int d = 0; d < T.mi[m].dim; ++d)
      if (B.level > 0 && (B.nx[d] & 1)) return "refined meshes need even block sizes on active dims";
    for (int d = T.mi[m].dim; d < 3; ++d)
      if (B.lx[d] != 0) return "logical location must be 0 on inactive dims";
    if (B.level > lmax[m]) lmax[m] = B.level;
    // no block may lie inside another
    for (int L = B.level - 1; L >= 0; --L) {
      const int sh = B.level - L;
      const Key k{m, L, B.lx[0] >> (T.mi[m].dim > 0 ? sh : 0), B.lx[1] >> (T.mi[m].dim > 1 ? sh : 0),
                  B.lx[2] >> (T.mi[m].dim > 2 ? sh : 0)};
      if (T.at.count(k)) return "block " + std::to_string(b) + " overlaps a coarser block";
    }
  }
  // and together they cover the domain: volumes in finest cells add up
  std::vector<long double> vol(nmesh, 0.0L);
  for (size_t b = 0; b < hb.size(); ++b) {
    const int m = bmesh[b];
    vol[m] += (long double)((int64_t)1 << ((lmax[m] - hb[b].level) * T.mi[m].dim));
  }
  for (int m = 0; m < nmesh; ++m) {
    long double want = 1.0L;
    for (int d = 0; d < T.mi[m].dim; ++d) want *= (long double)((int64_t)T.mi[m].nb[d] << lmax[m]);
    if (vol[m] != want) return "blocks of mesh " + std::to_string(m) + " do not tile its domain";
  }
  return "";
}

std::string build_refinement_tables(const std::vector<HostBlock> &hb, const std::vector<int> &bmesh,
                                    const std::vector<pangu_config> &mcfg, std::vector<XFill> &fill,
                                    std::vector<XProl> &prol, std::vector<XFlux> &flux) {
  Tree T(hb, bmesh, mcfg);
  if (!T.err.empty()) return T.err;
  fill.clear();
  prol.clear();
  flux.clear();
  for (int b = 0; b < (int)hb.size(); ++b) {
    const HostBlock &F = hb[b];
    const int m = bmesh[b], L = F.level, dim = T.mi[m].dim;
    std::map<std::array<int64_t, 3>, int> prol_of;  // coarse cell (unwrapped) -> entry
    for (int k = 0; k < F.n[2]; ++k)
      for (int j = 0; j < F.n[1]; ++j)
        for (int i = 0; i < F.n[0]; ++i) {
          const int ix[3] = {i, j, k};
          bool interior = true, outside = false;
          int64_t gu[3] = {0, 0, 0}, g[3] = {0, 0, 0};
          for (int d = 0; d < dim; ++d) {
            if (ix[d] < F.s[d] || ix[d] > F.e[d]) interior = false;
            gu[d] = (int64_t)F.lx[d] * F.nx[d] + (ix[d] - F.s[d]);
            g[d] = gu[d];
            if (!T.wrap(m, d, L, g[d])) outside = true;
          }
          if (interior || outside) continue;  // physical boundaries fill the rest
          int lev = 0;
          const int N = T.cover(m, L, g, lev);
          if (N < 0) return "neighbouring blocks differ by more than one level (block " + std::to_string(b) + ")";
          const int dst = (int)F.id(k, j, i);
          if (lev == L) {
            fill.push_back({b, dst, N, T.cell(hb[N], g), 0});
            continue;
          }
          if (lev == L + 1) {
            int64_t f[3];
            for (int d = 0; d < 3; ++d) f[d] = d < dim ? 2 * g[d] : 0;
            fill.push_back({b, dst, N, T.cell(hb[N], f), 1});
            continue;
          }
          // covered by a coarser block: prolongation of coarse cell G
          std::array<int64_t, 3> G = {0, 0, 0};
          for (int d = 0; d < dim; ++d) G[d] = floor2(gu[d]);
          auto it = prol_of.find(G);
          int e;
          if (it != prol_of.end()) {
            e = it->second;
          } else {
            XProl P;
            P.b = b;
            P.mask = 0;
            int lo[3] = {F.s[0], F.s[1], F.s[2]};
            for (int d = 0; d < dim; ++d) lo[d] = (int)(2 * G[d] - (int64_t)F.lx[d] * F.nx[d] + F.s[d]);
            P.base = (int)F.id(lo[2], lo[1], lo[0]);
            for (int s = 0; s < 7; ++s) {
              P.sb[s] = -1;
              P.sc[s] = 0;
            }
            for (int s = 0; s < 1 + 2 * dim; ++s) {
              int64_t X[3] = {G[0], G[1], G[2]};
              if (s > 0) X[(s - 1) / 2] += (s & 1) ? -1 : 1;
              int64_t Xw[3] = {X[0], X[1], X[2]};
              for (int d = 0; d < dim; ++d)
                if (!T.wrap(m, d, L - 1, Xw[d]))
                  return "prolongation stencil crosses a physical boundary (block " + std::to_string(b) + ")";
              const int C = T.find(m, L - 1, Xw);
              if (C >= 0) {
                P.sb[s] = C;
                P.sc[s] = T.cell(hb[C], Xw);
                continue;
              }
              if (s == 0) return "internal error: prolongated cell not coarse";
              // a coarse value where the mesh is at least as fine as this
              // block: the mean of this block's own 2^dim cells there
              int fl[3] = {F.s[0], F.s[1], F.s[2]};
              int64_t fw[3] = {0, 0, 0};
              for (int d = 0; d < dim; ++d) {
                fl[d] = (int)(2 * X[d] - (int64_t)F.lx[d] * F.nx[d] + F.s[d]);
                if (fl[d] < 0 || fl[d] + 1 > F.n[d] - 1)
                  return "prolongation stencil beyond the ghost zones (block " + std::to_string(b) + ")";
                fw[d] = 2 * Xw[d];
              }
              int lev2 = 0;
              if (T.cover(m, L, fw, lev2) < 0 || lev2 < L)
                return "prolongation stencil over a prolongated region (block " + std::to_string(b) + ")";
              P.sb[s] = -1;
              P.sc[s] = (int)F.id(fl[2], fl[1], fl[0]);
            }
            e = (int)prol.size();
            prol.push_back(P);
            prol_of[G] = e;
          }
          int c = 0;
          for (int d = 0; d < dim; ++d) c |= (int)(gu[d] - 2 * G[d]) << d;
          prol[e].mask |= 1 << c;
        }
    // flux correction: faces of b shared with finer blocks
    for (int d = 0; d < dim; ++d)
      for (int side = 0; side < 2; ++side) {
        const int face = side ? F.e[d] + 1 : F.s[d];
        const int out = side ? F.e[d] + 1 : F.s[d] - 1;
        int lo[3], hi[3];
        for (int t = 0; t < 3; ++t) {
          lo[t] = F.s[t];
          hi[t] = F.e[t];
        }
        lo[d] = hi[d] = out;
        for (int k = lo[2]; k <= hi[2]; ++k)
          for (int j = lo[1]; j <= hi[1]; ++j)
            for (int i = lo[0]; i <= hi[0]; ++i) {
              const int ix[3] = {i, j, k};
              int64_t g[3] = {0, 0, 0};
              bool outside = false;
              for (int t = 0; t < dim; ++t) {
                g[t] = (int64_t)F.lx[t] * F.nx[t] + (ix[t] - F.s[t]);
                if (!T.wrap(m, t, L, g[t])) outside = true;
              }
              if (outside) continue;
              int64_t f[3] = {0, 0, 0};
              for (int t = 0; t < dim; ++t) f[t] = 2 * g[t];
              const int N = T.find(m, L + 1, f);
              if (N < 0) continue;
              const HostBlock &NB = hb[N];
              int fl[3] = {NB.s[0], NB.s[1], NB.s[2]};
              for (int t = 0; t < dim; ++t) fl[t] = T.local(NB, t, f[t]);
              if (!side) fl[d] += 2;  // the upper face of fine cell 2g + 1
              int cl[3] = {i, j, k};
              cl[d] = face;
              flux.push_back({b, (int)F.id(cl[2], cl[1], cl[0]), N, (int)NB.id(fl[2], fl[1], fl[0]), d});
            }
      }
  }
  return "";
}

}  // namespace pangu
