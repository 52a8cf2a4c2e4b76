// Shared host/device definitions of pangu_hip.
#pragma once
#include <stdint.h>

#ifndef __HIPCC__
#define __host__
#define __device__
#endif

#include "../../include/pangu_hip.h"

namespace pangu {

#ifdef __HIPCC__
// Every data pointer the kernels receive points into device global memory,
// but a pointer loaded from a pointer table is generic to the compiler, and
// accesses through it become flat_* instructions (which also count against
// the LDS/scalar counter and take a 64-bit vector address).  Accesses
// through glb(p) compile to global_* instructions (scalar base + 32-bit
// vector offset where possible).
template <class T>
using gptr = __attribute__((address_space(1))) T *;
template <class T>
__device__ __forceinline__ gptr<T> glb(T *p) {
  return (gptr<T>)p;
}
// Element i of a device array through a 32-bit byte offset: with a base
// the compiler knows to be uniform this is one global_* instruction with
// the base in scalar registers and one 32-bit vector offset shared by
// neighbouring accesses, instead of a 64-bit address computed per access.
// The per-thread index times the element size must stay below 4 GiB
// (pangu_create rejects blocks whose geometry tables would exceed 2^23
// entries, which keeps every index the kernels form within that).
// ld32(base, u, i) reads base[u + i] where u is the same for the whole
// workgroup (a stencil offset, a table component) and i the thread's own
// index: base + u goes to scalar registers, i to the vector offset.
template <class T>
__device__ __forceinline__ T ld32(const T *base, int64_t u, uint32_t i) {
  return *(const __attribute__((address_space(1))) T *)((const __attribute__((address_space(1))) char *)(base + u) +
                                                        i * (uint32_t)sizeof(T));
}
// ld32 through the constant address space, for the geometry tables, which
// no kernel writes: the compiler may then move the load across the kernel's
// stores (which it must otherwise assume could alias the table).
template <class T>
__device__ __forceinline__ T ldc32(const T *base, int64_t u, uint32_t i) {
  return *(const __attribute__((address_space(4))) T *)((const __attribute__((address_space(4))) char *)(base + u) +
                                                        i * (uint32_t)sizeof(T));
}
template <class T>
__device__ __forceinline__ void st32(T *base, uint32_t i, T v) {
  *(__attribute__((address_space(1))) T *)((__attribute__((address_space(1))) char *)base +
                                           i * (uint32_t)sizeof(T)) = v;
}
// An array read and written element-wise through ld32 / st32: a[i] reads
// or assigns element i.
template <class T>
struct Ref32 {
  T *p;
  uint32_t i;
  __device__ __forceinline__ operator T() const { return ld32(p, 0, i); }
  __device__ __forceinline__ const Ref32 &operator=(T v) const {
    st32(p, i, v);
    return *this;
  }
};
template <class T>
struct Arr32 {
  T *p;
  __device__ __forceinline__ Ref32<T> operator[](int64_t i) const { return Ref32<T>{p, (uint32_t)i}; }
};
#endif

// Static mesh refinement: gather tables built on the host from the block
// tree (refine_host.cpp).  Cell indices are within a block's ghosted array.
struct XFill {   // a ghost cell filled from another block's interior
  int db, dst;   // destination block and cell
  int sb, src;   // source block; its cell (copy) or the lowest of the 2^dim fine cells (mean)
  int mean;      // 0 copy, 1 mean of the 2^dim fine cells a coarse ghost cell covers
};
struct XProl {   // a coarse cell prolongated into the fine ghost cells it covers
  int b, base;   // fine block and the lowest fine cell the coarse cell covers
  int mask;      // bit c0 + 2 c1 + 4 c2 set: child (c0, c1, c2) is a ghost cell of b
  int sb[7];     // coarse values centre, -x1, +x1, -x2, +x2, -x3, +x3: a coarse block,
  int sc[7];     // its cell; or sb = -1: the mean of b's own 2^dim fine cells from sc
};
struct XFlux {   // a coarse face flux replaced by the mean of the fine ones covering it
  int cb, cface; // coarse block, its face (flux index along d)
  int fb, fface; // fine block, the lowest of the 2^(dim-1) fine faces
  int d;
};

// Slot order of a full state vector (variable_mnemonics.h:7-10).
enum Slot { RHO = 0, ENY, UX1, UX2, UX3, BX1, BX2, BX3, ENT, KEL, NPRIM };

// Geometry table components.  Every derived quantity is computed on the host
// in FP64 with IEEE operations only, so device kernels consume the exact
// values a per-cell recomputation would produce.
//
// Cell-centre table (one entry per ghosted (j,i), or (k,j,i) for x3-dependent
// metrics):
enum CentreComp {
  CG_GCOV = 0,          // 16: gcov[r][c] at CG_GCOV + 4r + c
  CG_GCON = 16,         // 16: gcon[r][c]
  CG_SG = 32,           // sqrt|det g|
  CG_ALPHA = 33,        // 1/sqrt(-gcon00)   (state lapse)
  CG_LAPSE = 34,        // sqrt(-1/gcon00)   (recovery normal observer)
  CG_RHOFL = 35,        // density_floor * max(r,1)^density_floor_pow
  CG_ENGFL = 36,        // energy_floor  * max(r,1)^energy_floor_pow
  CG_RAD = 37,          // r used by the floors / excision
  CG_CONN = 38,         // 64: Gamma^a_{bc} at CG_CONN + 16a + 4b + c, or, when the
                        // set's conn_sym is set, 40 at CG_CONN + 10a + pair(b, c)
  CG_NCOMP = 102
};
// Face table (lower faces normal to direction d):
enum FaceComp {
  FG_GCOV = 0,    // 16
  FG_GCON0 = 16,  // 4: gcon[0][0..3]
  FG_GCONDD = 20, // gcon[d][d]
  FG_SG = 21,     // sqrt|det g|
  FG_ALPHA = 22,  // 1/sqrt(-gcon00)
  FG_VALID = 23,  // 1.0, or 0.0 where the metric is singular (polar axis)
  FG_NCOMP = 24
};

// Per-block launch descriptor (device copy).  Blocks of one context may
// differ in shape and belong to different meshes (independent simulations).
struct DevBlock {
  int n[3];     // ghosted extent (1 on inactive dims)
  int s[3];     // interior start
  int e[3];     // interior end (inclusive)
  int lx[3];    // logical location in its mesh, at its refinement level
  double dx[3];
  int64_t ncell;        // n0*n1*n2
  int geo;              // geometry set index
  int mesh;             // mesh the block belongs to
  int bcf[3][2];        // physical boundary kind (PANGU_BC_*) of each face, -1 if exchanged
  int level;            // refinement level
  int pad;
};

// Geometry set (device copy): SoA tables, comp-major, (k?) j, i fastest.
struct DevGeo {
  const double *cen;    // CG_NCOMP * ncen
  const double *face[3];// FG_NCOMP * nface[d]; face[2] may alias the centre set
  int ni, nj, nk;       // table extent (nk = 1 for x3-independent metrics)
  int has3d;
  int64_t ncen;         // ni*nj*nk
  int64_t nface[3];
  int fi[3], fj[3];     // face table row lengths (ni+1 / nj+1 on the normal dim)
  int conn_sym;         // every Gamma^a_{bc} equals Gamma^a_{cb} bit for bit: compact storage
  int pad;
};

// Index of the unordered pair (b, c) among (0,0) (0,1) (0,2) (0,3) (1,1) (1,2)
// (1,3) (2,2) (2,3) (3,3).
__host__ __device__ constexpr int conn_pair(int b, int c) {
  return b <= c ? (b * (7 - b)) / 2 + c : (c * (7 - c)) / 2 + b;
}

}  // namespace pangu
