// Thread -> (block, k, j, i) decoding of tile launches (see Range3, Items).
#pragma once
#include <hip/hip_runtime.h>

#include "common.h"
#include "kernels.h"

namespace pangu {

constexpr int TILE = 256;

// Range of a launch on one block (see Range3).
__device__ __forceinline__ void block_range(const DevBlock &B, const Range3 &R, int lo[3],
                                            int cnt[3]) {
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    int l, h;
    if (R.entire) {
      l = 0;
      h = B.n[d] - 1;
    } else if (B.n[d] > 1) {
      l = B.s[d] + R.rlo[d];
      h = B.e[d] + R.rhi[d];
    } else {
      l = B.s[d];
      h = B.e[d];
    }
    lo[d] = l;
    cnt[d] = h - l + 1;
  }
}

// Decodes (block, k, j, i) of this thread; false for padding threads.  The
// workgroup index L is remapped to tile T = (L mod 8) * ceil(N/8) + L / 8,
// so each XCD runs a contiguous run of tiles.  Tiles hold R.tile cells of a
// plane (the workgroup size of the launch).
__device__ __forceinline__ bool decode_tile_at(const KCtx &C, const Range3 &R, int T, int &b, int &k,
                                               int &j, int &i);

__device__ __forceinline__ bool decode_tile(const KCtx &C, const Range3 &R, int &b, int &k, int &j,
                                            int &i) {
  const int N = R.ntiles;
  const int n8 = (N + 7) >> 3;
  const int L = (int)blockIdx.x;
  const int T = (L & 7) * n8 + (L >> 3);
  if (T >= N) return false;
  return decode_tile_at(C, R, T, b, k, j, i);
}

// (block, k, j, i) of this thread in tile T (0 <= T < R.ntiles).
__device__ __forceinline__ bool decode_tile_at(const KCtx &C, const Range3 &R, int T, int &b, int &k,
                                               int &j, int &i) {
  const int TL = R.tile;  // cells per tile, the workgroup size
  int lo[3], cnt[3], local, tpp;
  if (R.uni) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      lo[d] = R.ulo[d];
      cnt[d] = R.ucnt[d];
    }
    tpp = R.utpp;
    const int per_block = cnt[2] * tpp;
    b = T / per_block;
    local = T - b * per_block;
  } else {
    b = R.tblock[T];
    local = T - R.tfirst[b];
    block_range(C.blk[b], R, lo, cnt);
    tpp = (cnt[0] * cnt[1] + TL - 1) / TL;
  }
  const int plane = cnt[0] * cnt[1];
  int tile, kk;
  if (R.order) {
    tile = local / cnt[2];
    kk = local - tile * cnt[2];
  } else {
    kk = local / tpp;
    tile = local - kk * tpp;
  }
  const int t = tile * TL + (int)threadIdx.x;
  if (t >= plane) return false;
  k = lo[2] + kk;
  const int jj = t / cnt[0];
  j = lo[1] + jj;
  i = lo[0] + (t - jj * cnt[0]);
  return true;
}

// Lane-pair launches: every 256-cell tile of R is split over two
// workgroups of 128 lane pairs; thread t handles cell (t >> 1) of its half
// tile, and the two lanes of a pair share it.  Grid: 2 * ntiles workgroups
// (padded to a multiple of 8), remapped to tiles like decode_tile.
__device__ __forceinline__ bool decode_tile_pair(const KCtx &C, const Range3 &R, int &b, int &k, int &j,
                                                 int &i) {
  const int N2 = 2 * R.ntiles;
  const int n8 = (N2 + 7) >> 3;
  const int L = (int)blockIdx.x;
  const int T2 = (L & 7) * n8 + (L >> 3);
  if (T2 >= N2) return false;
  const int T = T2 >> 1, sub = T2 & 1;
  int lo[3], cnt[3], local, tpp;
  if (R.uni) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      lo[d] = R.ulo[d];
      cnt[d] = R.ucnt[d];
    }
    tpp = R.utpp;
    const int per_block = cnt[2] * tpp;
    b = T / per_block;
    local = T - b * per_block;
  } else {
    b = R.tblock[T];
    local = T - R.tfirst[b];
    block_range(C.blk[b], R, lo, cnt);
    tpp = (cnt[0] * cnt[1] + TILE - 1) / TILE;
  }
  const int plane = cnt[0] * cnt[1];
  int tile, kk;
  if (R.order) {
    tile = local / cnt[2];
    kk = local - tile * cnt[2];
  } else {
    kk = local / tpp;
    tile = local - kk * tpp;
  }
  const int t = tile * TILE + sub * (TILE / 2) + ((int)threadIdx.x >> 1);
  if (t >= plane) return false;
  k = lo[2] + kk;
  const int jj = t / cnt[0];
  j = lo[1] + jj;
  i = lo[0] + (t - jj * cnt[0]);
  return true;
}

// Marching launches of the sweep along DIR (2 or 3): a thread owns one
// position of the plane spanned by the other two directions (x1 fastest,
// 256 positions per tile) and R.step consecutive faces along DIR from
// `m` on (m < mend).  Tiles per block: ceil(count along DIR / step) chunks
// of ceil(plane / 256) tiles, ordered chunk-major (R.order 0) or
// tile-major (1), and remapped to XCDs like decode_tile.
template <int DIR>
__device__ __forceinline__ bool decode_march(const KCtx &C, const Range3 &R, int &b, int &k, int &j,
                                             int &i, int &mend) {
  const int N = R.ntiles;
  const int n8 = (N + 7) >> 3;
  const int L = (int)blockIdx.x;
  const int T = (L & 7) * n8 + (L >> 3);
  if (T >= N) return false;
  constexpr int D = DIR - 1, E = (DIR == 2) ? 2 : 1;  // marched and transverse dims
  int lo[3], cnt[3], local;
  if (R.uni) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      lo[d] = R.ulo[d];
      cnt[d] = R.ucnt[d];
    }
  } else {
    b = R.tblock[T];
    block_range(C.blk[b], R, lo, cnt);
  }
  const int plane = cnt[0] * cnt[E];
  const int tpp = (plane + TILE - 1) / TILE;
  const int nch = (cnt[D] + R.step - 1) / R.step;
  if (R.uni) {
    const int per_block = nch * tpp;
    b = T / per_block;
    local = T - b * per_block;
  } else {
    local = T - R.tfirst[b];
  }
  int tile, ch;
  if (R.order) {
    tile = local / nch;
    ch = local - tile * nch;
  } else {
    ch = local / tpp;
    tile = local - ch * tpp;
  }
  const int t = tile * TILE + (int)threadIdx.x;
  if (t >= plane) return false;
  const int u = t / cnt[0];
  i = lo[0] + (t - u * cnt[0]);
  const int m0 = lo[D] + ch * R.step;
  const int mhi = lo[D] + cnt[D];
  mend = (m0 + R.step < mhi) ? m0 + R.step : mhi;
  if (DIR == 2) {
    k = lo[2] + u;
    j = m0;
  } else {
    j = lo[1] + u;
    k = m0;
  }
  return true;
}

// Block and item index of this thread in an item launch.
__device__ __forceinline__ void decode_item(const Items &I, int &b, int &item) {
  b = I.tblock[blockIdx.x];
  item = ((int)blockIdx.x - I.tfirst[b]) * TILE + (int)threadIdx.x;
}

}  // namespace pangu
