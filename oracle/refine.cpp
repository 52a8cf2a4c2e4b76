// Static mesh refinement in the oracle.  TEST INFRASTRUCTURE ONLY.
//
// On a refined mesh PANGU's task list relies on Parthenon for three things
// (ideal_grmhd.cc:46-47,71-72,78): the block tree of <parthenon/static_
// refinementN> regions (fm_torus_cks/inputfile:87-122), ghost cells across
// refinement levels and flux correction.  Parthenon is a git submodule that
// is empty in the reference checkout, so these follow its published
// algorithm (Grete et al. 2022, IJHPCA 37:465; the Athena++ mesh-refinement
// scheme it inherits, Stone et al. 2020, ApJS 249:4), restated for a uniform
// Cartesian code-coordinate grid:
//  - leaves: per active dim at level L (nb << L blocks), a region spans the
//    blocks from the first whose upper edge exceeds its lower bound to the
//    first whose upper edge reaches its upper bound, widened to sibling
//    pairs; a block is refined only after its neighbours (faces, edges,
//    corners) are at least as fine, so adjacent leaves differ by <= 1 level;
//  - restriction: the mean of the 2^dim fine cells (equal code volumes);
//  - prolongation: the coarse value plus, per active dim, +-1/4 of the
//    minmod of the differences to the two coarse neighbours (the
//    limited-linear operator at equal spacing);
//  - flux correction: a coarse face flux shared with finer blocks becomes the
//    mean of the 2^(dim-1) fine face fluxes (equal face areas).
// The summation orders are fixed (x1 pairs first, then x2, then x3); the
// recovery flag restricts by minimum and prolongates by copy.
// Pinned by tests/test_refinement.py (hand-counted trees, exact exchange of
// linear data, an everywhere-refined mesh equal to the uniform mesh, mass
// conservation across levels).  Parity with Parthenon's own code is unpinned:
// that code is absent from the reference checkout.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <functional>
#include <set>

#include "oracle_sim.h"

namespace oracle {

static int64_t floor_half(int64_t g) { return (g - (g & 1)) / 2; }

void Sim::make_leaves(std::vector<std::array<int64_t, 4>> &out) const {
  const int dim = ndim;
  const int64_t nb[3] = {cfg.nb[0], cfg.nb[1], cfg.nb[2]};
  std::set<std::array<int64_t, 4>> S;
  for (int64_t k = 0; k < nb[2]; ++k)
    for (int64_t j = 0; j < nb[1]; ++j)
      for (int64_t i = 0; i < nb[0]; ++i) S.insert({0, i, j, k});
  // the leaf at or above (level, l); level -1 if (level, l) is refined further
  auto covering = [&](int64_t level, const int64_t l[3]) {
    for (int64_t L = level; L >= 0; --L) {
      std::array<int64_t, 4> key = {L, 0, 0, 0};
      for (int d = 0; d < dim; ++d) key[1 + d] = l[d] >> (level - L);
      if (S.count(key)) return key;
    }
    return std::array<int64_t, 4>{-1, 0, 0, 0};
  };
  std::function<void(int64_t, const int64_t *)> make_node;
  std::function<void(std::array<int64_t, 4>)> split = [&](std::array<int64_t, 4> leaf) {
    const int64_t L = leaf[0];
    const int r2 = dim > 2 ? 1 : 0, r1 = dim > 1 ? 1 : 0;
    for (int o2 = -r2; o2 <= r2; ++o2)
      for (int o1 = -r1; o1 <= r1; ++o1)
        for (int o0 = -1; o0 <= 1; ++o0) {
          const int o[3] = {o0, o1, o2};
          int64_t nl[3] = {0, 0, 0};
          bool in = true;
          for (int d = 0; d < dim; ++d) {
            const int64_t N = nb[d] << L;
            nl[d] = leaf[1 + d] + o[d];
            if (nl[d] < 0 || nl[d] >= N) {
              if (cfg.bc[d][0] == BC_PERIODIC)
                nl[d] = (nl[d] + N) % N;
              else
                in = false;
            }
          }
          if (!in) continue;
          const auto c = covering(L, nl);
          if (c[0] >= 0 && c[0] < L) make_node(L, nl);
        }
    if (!S.count(leaf)) return;
    S.erase(leaf);
    for (int c = 0; c < (1 << dim); ++c) {
      std::array<int64_t, 4> ch = {L + 1, 0, 0, 0};
      for (int d = 0; d < dim; ++d) ch[1 + d] = 2 * leaf[1 + d] + ((c >> d) & 1);
      S.insert(ch);
    }
  };
  make_node = [&](int64_t L, const int64_t *l) {
    for (;;) {
      const auto c = covering(L, l);
      if (c[0] < 0 || c[0] == L) return;
      split(c);
    }
  };
  for (int r = 0; r < cfg.nref; ++r) {
    const int L = cfg.ref_level[r];
    if (L <= 0) continue;
    int64_t lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    bool empty = false;
    for (int d = 0; d < dim; ++d) {
      const int64_t N = nb[d] << L;
      int64_t a = 0;
      while (a < N && !(mesh_edge(a + 1, N, cfg.xmin[d], cfg.xmax[d]) > cfg.ref_min[r][d])) ++a;
      int64_t b = a;
      while (b < N && !(mesh_edge(b + 1, N, cfg.xmin[d], cfg.xmax[d]) >= cfg.ref_max[r][d])) ++b;
      if (a >= N) empty = true;
      if (b >= N) b = N - 1;
      if (a & 1) --a;
      if (!(b & 1)) ++b;
      lo[d] = a;
      hi[d] = b;
    }
    if (empty) continue;
    for (int64_t k = lo[2]; k <= hi[2]; k += (dim > 2 ? 2 : 1))
      for (int64_t j = lo[1]; j <= hi[1]; j += (dim > 1 ? 2 : 1))
        for (int64_t i = lo[0]; i <= hi[0]; i += 2) {
          const int64_t l[3] = {i, j, k};
          make_node(L, l);
        }
  }
  // depth-first order: root blocks x1 fastest, then children x1 fastest
  std::vector<std::pair<std::vector<int64_t>, std::array<int64_t, 4>>> keyed;
  for (const auto &lf : S) {
    const int64_t L = lf[0];
    std::vector<int64_t> path;
    int64_t root = 0;
    for (int d = dim - 1; d >= 0; --d) root = root * nb[d] + (lf[1 + d] >> L);
    path.push_back(root);
    for (int64_t i = 1; i <= L; ++i) {
      int64_t c = 0;
      for (int d = 0; d < dim; ++d) c |= ((lf[1 + d] >> (L - i)) & 1) << d;
      path.push_back(c);
    }
    keyed.push_back({path, lf});
  }
  std::sort(keyed.begin(), keyed.end());
  out.clear();
  for (const auto &kv : keyed) out.push_back(kv.second);
}

bool Sim::wrap_cell(int d, int level, int64_t &g) const {
  if (d >= ndim) {
    g = 0;
    return true;
  }
  const int64_t N = ((int64_t)cfg.nb[d] * blk[0].nx[d]) << level;
  if (g >= 0 && g < N) return true;
  if (cfg.bc[d][0] != BC_PERIODIC) return false;
  g = ((g % N) + N) % N;
  return true;
}

int Sim::leaf_at(int level, const int64_t g[3]) const {
  if (level < 0) return -1;
  std::array<int64_t, 4> key = {level, 0, 0, 0};
  for (int d = 0; d < ndim; ++d) key[1 + d] = g[d] / blk[0].nx[d];
  auto it = leaf_of.find(key);
  return it == leaf_of.end() ? -1 : it->second;
}

static int64_t local_id(const Block &B, int ndim, const int64_t g[3]) {
  int ix[3] = {B.s[0], B.s[1], B.s[2]};
  for (int d = 0; d < ndim; ++d) ix[d] = (int)(g[d] - (int64_t)B.lx[d] * B.nx[d] + B.s[d]);
  return B.idx(ix[2], ix[1], ix[0]);
}

// Mean (q >= 0: primitive q) or minimum (q < 0: flag) of the 2^dim cells of
// block b from cell g (level of b, x1 fastest), in the fixed pair order.
static double fine_reduce(const Sim &s, int c, int q, int b, const int64_t g[3]) {
  const Block &B = s.blk[b];
  const double *a = q >= 0 ? s.Pc(c, b, q) : s.flag.data() + B.off;
  const int64_t i0 = local_id(B, s.ndim, g), sj = B.n[0], sk = (int64_t)B.n[0] * B.n[1];
  if (q < 0) {
    double m = std::fmin(a[i0], a[i0 + 1]);
    if (s.ndim == 1) return m;
    m = std::fmin(m, std::fmin(a[i0 + sj], a[i0 + sj + 1]));
    if (s.ndim == 2) return m;
    return std::fmin(m, std::fmin(std::fmin(a[i0 + sk], a[i0 + sk + 1]),
                                  std::fmin(a[i0 + sk + sj], a[i0 + sk + sj + 1])));
  }
  if (s.ndim == 1) return (a[i0] + a[i0 + 1]) * 0.5;
  const double lo = (a[i0] + a[i0 + 1]) + (a[i0 + sj] + a[i0 + sj + 1]);
  if (s.ndim == 2) return lo * 0.25;
  const double hi = (a[i0 + sk] + a[i0 + sk + 1]) + (a[i0 + sk + sj] + a[i0 + sk + sj + 1]);
  return (lo + hi) * 0.125;
}

// The mesh's value at level-`level` cell g (wrapped): a leaf's cell at that
// level, or the restriction of the leaf one level finer.
double Sim::level_value(int c, int q, int level, const int64_t g[3]) const {
  const int b = leaf_at(level, g);
  if (b >= 0) {
    const int64_t id = local_id(blk[b], ndim, g);
    return q >= 0 ? Pc(c, b, q)[id] : flag[blk[b].off + id];
  }
  int64_t f[3] = {0, 0, 0};
  for (int d = 0; d < ndim; ++d) f[d] = 2 * g[d];
  const int bf = leaf_at(level + 1, f);
  if (bf < 0) {
    fprintf(stderr, "oracle: no leaf within one level of a stencil cell\n");
    return NAN;
  }
  return fine_reduce(*this, c, q, bf, f);
}

static double minmod(double a, double b) {
  if (a > 0.0 && b > 0.0) return a < b ? a : b;
  if (a < 0.0 && b < 0.0) return a > b ? a : b;
  return 0.0;
}

void Sim::exchange_refined(int c) {
  for (int b = 0; b < nblk; ++b) {
    const Block &B = blk[b];
    const int L = B.level;
    for (int k = 0; k < B.n[2]; ++k)
      for (int j = 0; j < B.n[1]; ++j)
        for (int i = 0; i < B.n[0]; ++i) {
          const int ix[3] = {i, j, k};
          bool interior = true, outside = false;
          int64_t gu[3] = {0, 0, 0}, g[3] = {0, 0, 0};
          for (int d = 0; d < ndim; ++d) {
            if (ix[d] < B.s[d] || ix[d] > B.e[d]) interior = false;
            gu[d] = (int64_t)B.lx[d] * B.nx[d] + (ix[d] - B.s[d]);
            g[d] = gu[d];
            if (!wrap_cell(d, L, g[d])) outside = true;
          }
          if (interior || outside) continue;
          const int64_t dst = B.idx(k, j, i);
          int64_t G[3] = {0, 0, 0};
          for (int d = 0; d < ndim; ++d) G[d] = g[d] >> 1;
          const int bc = leaf_at(L - 1, G);
          if (bc < 0) {
            // a leaf of this level or one level finer covers the cell
            for (int q = 0; q < np; ++q) Pc(c, b, q)[dst] = level_value(c, q, L, g);
            flag[B.off + dst] = level_value(c, -1, L, g);
            continue;
          }
          // prolongation from coarse cell Gu (unwrapped) of leaf bc
          int64_t Gu[3] = {0, 0, 0};
          for (int d = 0; d < ndim; ++d) Gu[d] = floor_half(gu[d]);
          auto coarse = [&](int q, int d, int side) {
            int64_t X[3] = {Gu[0], Gu[1], Gu[2]};
            if (d >= 0) X[d] += side;
            for (int e = 0; e < ndim; ++e)
              if (!wrap_cell(e, L - 1, X[e])) {
                fprintf(stderr, "oracle: prolongation stencil crosses a physical boundary\n");
                return (double)NAN;
              }
            const int bb = leaf_at(L - 1, X);
            if (bb >= 0) {
              const int64_t id = local_id(blk[bb], ndim, X);
              return q >= 0 ? Pc(c, bb, q)[id] : flag[blk[bb].off + id];
            }
            // finer than the coarse level here: restrict the level-L values
            double v[8];
            for (int ch = 0; ch < (1 << ndim); ++ch) {
              int64_t f[3] = {0, 0, 0};
              for (int e = 0; e < ndim; ++e) f[e] = 2 * X[e] + ((ch >> e) & 1);
              v[ch] = level_value(c, q, L, f);
            }
            if (ndim == 1) return (v[0] + v[1]) * 0.5;
            const double lo = (v[0] + v[1]) + (v[2] + v[3]);
            if (ndim == 2) return lo * 0.25;
            return (lo + ((v[4] + v[5]) + (v[6] + v[7]))) * 0.125;
          };
          int child = 0;
          for (int d = 0; d < ndim; ++d) child |= (int)(gu[d] - 2 * Gu[d]) << d;
          for (int q = 0; q < np; ++q) {
            const double v0 = coarse(q, -1, 0);
            double w = v0;
            for (int d = 0; d < ndim; ++d) {
              const double sl = (g_plant == PLANT_PROLONG_SLOPE ? 0.5 : 0.25) *
                                minmod(v0 - coarse(q, d, -1), coarse(q, d, 1) - v0);
              w = ((child >> d) & 1) ? w + sl : w - sl;
            }
            Pc(c, b, q)[dst] = w;
          }
          flag[B.off + dst] = coarse(-1, -1, 0);
        }
  }
}

void Sim::flux_correction() {
  if (!refined || g_plant == PLANT_NO_FLUX_CORRECTION) return;
  for (int b = 0; b < nblk; ++b) {
    const Block &B = blk[b];
    const int L = B.level;
    for (int d = 0; d < ndim; ++d)
      for (int side = 0; side < 2; ++side) {
        const int face = side ? B.e[d] + 1 : B.s[d];
        const int out = side ? B.e[d] + 1 : B.s[d] - 1;
        int lo[3], hi[3];
        for (int t = 0; t < 3; ++t) {
          lo[t] = B.s[t];
          hi[t] = B.e[t];
        }
        lo[d] = hi[d] = out;
        for (int k = lo[2]; k <= hi[2]; ++k)
          for (int j = lo[1]; j <= hi[1]; ++j)
            for (int i = lo[0]; i <= hi[0]; ++i) {
              const int ix[3] = {i, j, k};
              int64_t g[3] = {0, 0, 0};
              bool outside = false;
              for (int t = 0; t < ndim; ++t) {
                g[t] = (int64_t)B.lx[t] * B.nx[t] + (ix[t] - B.s[t]);
                if (!wrap_cell(t, L, g[t])) outside = true;
              }
              if (outside) continue;
              int64_t f[3] = {0, 0, 0};
              for (int t = 0; t < ndim; ++t) f[t] = 2 * g[t];
              const int fb = leaf_at(L + 1, f);
              if (fb < 0) continue;
              const Block &N = blk[fb];
              int fl[3] = {N.s[0], N.s[1], N.s[2]};
              for (int t = 0; t < ndim; ++t) fl[t] = (int)(f[t] - (int64_t)N.lx[t] * N.nx[t] + N.s[t]);
              if (!side) fl[d] += 2;
              const int st[3] = {1, N.n[0], N.n[0] * N.n[1]};
              const int ta = st[d == 0 ? 1 : 0], tb = st[d == 2 ? 1 : 2];
              const int64_t fi = N.idx(fl[2], fl[1], fl[0]);
              int cl[3] = {i, j, k};
              cl[d] = face;
              const int64_t ci = B.idx(cl[2], cl[1], cl[0]);
              for (int q = 0; q < nc; ++q) {
                const double *ff = Fc(d, fb, q);
                double m;
                if (ndim == 1)
                  m = ff[fi];
                else if (ndim == 2)
                  m = (ff[fi] + ff[fi + ta]) * 0.5;
                else
                  m = ((ff[fi] + ff[fi + ta]) + (ff[fi + tb] + ff[fi + ta + tb])) * 0.25;
                Fc(d, b, q)[ci] = m;
              }
            }
      }
  }
}

}  // namespace oracle
