// CPU check of the refinement gather tables (pangu_amd/csrc/refine_host.cpp):
// every index the kernels of refine.hip would touch lies inside its block,
// and the kernels' arithmetic replayed on the host over the tables turns the
// oracle's state before exchange / flux correction into the oracle's state
// after them, bit for bit.  Built and loaded by tests/test_refine_tables.py.
#include <cstdio>
#include <cstring>
#include <vector>

#include "common.h"
#include "host.h"

using namespace pangu;

namespace {

double fine_mean(const double *a, long c, int dim, long sj, long sk) {
  if (dim == 1) return (a[c] + a[c + 1]) * 0.5;
  const double lo = (a[c] + a[c + 1]) + (a[c + sj] + a[c + sj + 1]);
  if (dim == 2) return lo * 0.25;
  const double hi = (a[c + sk] + a[c + sk + 1]) + (a[c + sk + sj] + a[c + sk + sj + 1]);
  return (lo + hi) * 0.125;
}

double minmod(double a, double b) {
  if (a > 0.0 && b > 0.0) return a < b ? a : b;
  if (a < 0.0 && b < 0.0) return a > b ? a : b;
  return 0.0;
}

}  // namespace

// blocks: per block {nx[3], ng[3], lx[3], level} (ints) and {xmin[3], dx[3]}
// (doubles); one mesh with root blocks nb[3] and boundaries bc[3][2].
// prim: [block][np][cells] before the exchange (replayed in place);
// want: the same after the oracle's exchange.  flux[d]: [block][nc][cells]
// before flux correction (replayed in place), want_flux[d] after.
// Returns 0, or 1 tiling rejected, 2 tables rejected, 3 an index out of
// bounds, 4 a replayed value differs.
extern "C" int check_tables(int nblk, const int *ib, const double *db, const int *nb, const int *bc,
                            int np, double *prim, const double *want, int nc, double *f0, double *f1,
                            double *f2, const double *wf0, const double *wf1, const double *wf2,
                            long *counts /* fill, prol, flux */) {
  std::vector<HostBlock> hb;
  std::vector<int> bmesh(nblk, 0);
  for (int b = 0; b < nblk; ++b) {
    pangu_block B;
    std::memset(&B, 0, sizeof(B));
    for (int d = 0; d < 3; ++d) {
      B.nx[d] = ib[b * 10 + d];
      B.ng[d] = ib[b * 10 + 3 + d];
      B.lx[d] = ib[b * 10 + 6 + d];
      B.xmin[d] = db[b * 6 + d];
      B.dx[d] = db[b * 6 + 3 + d];
    }
    B.level = ib[b * 10 + 9];
    hb.push_back(make_host_block(B));
  }
  std::vector<pangu_config> mc(1);
  std::memset(&mc[0], 0, sizeof(pangu_config));
  for (int d = 0; d < 3; ++d) {
    mc[0].nb[d] = nb[d];
    mc[0].bc[d][0] = bc[2 * d];
    mc[0].bc[d][1] = bc[2 * d + 1];
  }
  std::string why = check_tiling(hb, bmesh, mc);
  if (!why.empty()) {
    fprintf(stderr, "tiling: %s\n", why.c_str());
    return 1;
  }
  std::vector<XFill> fill;
  std::vector<XProl> prol;
  std::vector<XFlux> flux;
  why = build_refinement_tables(hb, bmesh, mc, fill, prol, flux);
  if (!why.empty()) {
    fprintf(stderr, "tables: %s\n", why.c_str());
    return 2;
  }
  counts[0] = (long)fill.size();
  counts[1] = (long)prol.size();
  counts[2] = (long)flux.size();

  const long cells = hb[0].cells;
  const int dim = hb[0].n[2] > 1 ? 3 : (hb[0].n[1] > 1 ? 2 : 1);
  const long sj = hb[0].n[0], sk = (long)hb[0].n[0] * hb[0].n[1];
  const long span = 1 + (dim > 1 ? sj : 0) + (dim > 2 ? sk : 0);  // last cell of a 2^dim group
  auto inside = [&](int b, long c) { return b >= 0 && b < nblk && c >= 0 && c < cells; };
  long bad = 0;
  for (const XFill &X : fill) {
    if (!inside(X.db, X.dst) || !inside(X.sb, X.src)) ++bad;
    if (X.mean && !inside(X.sb, X.src + span)) ++bad;
  }
  for (const XProl &X : prol) {
    for (int c = 0; c < (1 << dim); ++c)
      if ((X.mask >> c) & 1) {
        const long o = X.base + (c & 1) + ((c >> 1) & 1) * sj + ((c >> 2) & 1) * sk;
        if (!inside(X.b, o)) ++bad;
      }
    if (X.sb[0] < 0) ++bad;  // the centre is always a coarse cell
    for (int s = 0; s < 1 + 2 * dim; ++s) {
      if (X.sb[s] >= 0) {
        if (!inside(X.sb[s], X.sc[s])) ++bad;
      } else if (!inside(X.b, X.sc[s]) || !inside(X.b, X.sc[s] + span)) {
        ++bad;
      }
    }
  }
  const int stride[3] = {1, (int)sj, (int)sk};
  for (const XFlux &X : flux) {
    const int ta = stride[X.d == 0 ? 1 : 0], tb = stride[X.d == 2 ? 1 : 2];
    const long last = X.fface + (dim > 1 ? ta : 0) + (dim > 2 ? tb : 0);
    if (!inside(X.cb, X.cface) || !inside(X.fb, X.fface) || !inside(X.fb, last)) ++bad;
  }
  if (bad) {
    fprintf(stderr, "indices out of bounds: %ld\n", bad);
    return 3;
  }

  // replay k_xfill, then k_xprol, on prim
  auto P = [&](int b, int q) { return prim + ((long)b * np + q) * cells; };
  for (const XFill &X : fill)
    for (int q = 0; q < np; ++q)
      P(X.db, q)[X.dst] = X.mean ? fine_mean(P(X.sb, q), X.src, dim, sj, sk) : P(X.sb, q)[X.src];
  for (const XProl &X : prol)
    for (int q = 0; q < np; ++q) {
      double v[7];
      for (int s = 0; s < 1 + 2 * dim; ++s)
        v[s] = X.sb[s] >= 0 ? P(X.sb[s], q)[X.sc[s]] : fine_mean(P(X.b, q), X.sc[s], dim, sj, sk);
      double sl[3];
      for (int d = 0; d < dim; ++d) sl[d] = 0.25 * minmod(v[0] - v[1 + 2 * d], v[2 + 2 * d] - v[0]);
      for (int c = 0; c < (1 << dim); ++c) {
        if (!((X.mask >> c) & 1)) continue;
        double w = v[0];
        for (int d = 0; d < dim; ++d) w = ((c >> d) & 1) ? w + sl[d] : w - sl[d];
        P(X.b, q)[X.base + (c & 1) + ((c >> 1) & 1) * sj + ((c >> 2) & 1) * sk] = w;
      }
    }
  long diff = 0;
  for (long n = 0; n < (long)nblk * np * cells; ++n)
    if (std::memcmp(&prim[n], &want[n], sizeof(double)) != 0) ++diff;

  // replay k_flux_correct
  double *F[3] = {f0, f1, f2};
  const double *WF[3] = {wf0, wf1, wf2};
  for (const XFlux &X : flux) {
    const int ta = stride[X.d == 0 ? 1 : 0], tb = stride[X.d == 2 ? 1 : 2];
    for (int q = 0; q < nc; ++q) {
      const double *f = F[X.d] + ((long)X.fb * nc + q) * cells;
      const long c = X.fface;
      const double m = dim == 1   ? f[c]
                       : dim == 2 ? (f[c] + f[c + ta]) * 0.5
                                  : ((f[c] + f[c + ta]) + (f[c + tb] + f[c + ta + tb])) * 0.25;
      F[X.d][((long)X.cb * nc + q) * cells + X.cface] = m;
    }
  }
  for (int d = 0; d < dim; ++d)
    for (long n = 0; n < (long)nblk * nc * cells; ++n)
      if (std::memcmp(&F[d][n], &WF[d][n], sizeof(double)) != 0) ++diff;
  if (diff) {
    fprintf(stderr, "replayed values differ: %ld\n", diff);
    return 4;
  }
  return 0;
}
