/*
 * pangu_phdf: Parthenon-style HDF5 output of cell fields (no counterpart in
 * the path itself: PANGU writes these files through Parthenon's output
 * machinery, <parthenon/output0> file_type = hdf5, write_xdmf = true in
 * fm_torus_mks/inputfile:66-70).
 *
 * The file holds what the reference's plotting scripts read
 * (scripts/python/xzplot.py:33-80, xz_temperature_plot.py:54-62):
 *   /<variable>              FP64 (nb, nz, ny, nx) for one component,
 *                            (nb, ncomp, nz, ny, nx) for several
 *   /VolumeLocations/{x,y,z} cell centres per block, (nb, n)
 *   /Locations/{x,y,z}       cell faces per block, (nb, n + 1)
 *   /LogicalLocations        int64 (nb, 3), each at its block's level;
 *                            /Blocks/loc.lx123 the same
 *   /Levels                  int32 (nb), refinement level over the root grid
 *   /Info                    attributes Time, NCycle, dt, NumDims,
 *                            NumMeshBlocks, MeshBlockSize, RootGridSize,
 *                            RootGridDomain, IncludesGhost (0), NGhost,
 *                            Coordinates, OutputDatasetNames, NumComponents
 * plus an XDMF sidecar <path>.xdmf (one rectilinear grid per block, cell
 * attributes as hyperslabs of the datasets) for VisIt / ParaView.
 * Host arrays only: the caller copies the fields off the GPU first.
 */
#ifndef PANGU_PHDF_H
#define PANGU_PHDF_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pangu_phdf_mesh {
  int nblocks;
  int nx[3];               /* interior cells per block and dim (1 = inactive) */
  int ng;                  /* ghost layers of the run (recorded, not written) */
  int ndim;
  int ncycle;
  double time, dt;
  int root_nx[3];          /* mesh cells per dim */
  double root_min[3], root_max[3];
  const int64_t *lx;       /* [nblocks][3] logical block locations */
  const double *xmin;      /* [nblocks][3] lower interior corner of each block */
  const double *dx;        /* [nblocks][3] cell widths */
  const int *level;        /* [nblocks] refinement levels, or NULL: all 0 */
} pangu_phdf_mesh;

typedef struct pangu_phdf_var {
  const char *name;
  int ncomp;
  const double *data;      /* [nblocks][ncomp][nz][ny][nx], interior cells */
} pangu_phdf_var;

/* Writes path (HDF5) and path.xdmf.  Returns 0, or -1 with a message that
 * pangu_phdf_last_error() returns. */
int pangu_write_phdf(const char *path, const pangu_phdf_mesh *mesh, int nvars,
                     const pangu_phdf_var *vars);
const char *pangu_phdf_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* PANGU_PHDF_H */
