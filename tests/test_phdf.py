"""Parthenon-style HDF5 output (include/pangu_phdf.h, pangu_amd/phdf.py).

Files are read back with h5dump (the image has no h5py).  The checks follow
what the reference's plotting script reads (scripts/python/xzplot.py:33-80):
per-variable datasets (nb, nz, ny, nx), VolumeLocations, LogicalLocations
and Info/Time; a numpy restatement of that script's block assembly must give
back the global field."""
import ctypes as C
import os
import re
import subprocess
import xml.etree.ElementTree as ET

import numpy as np
import pytest

H5DUMP = "/opt/conda/bin/h5dump"
pytestmark = pytest.mark.skipif(not os.path.exists(H5DUMP), reason="h5dump not present")


def _phdf():
    from pangu_amd import phdf
    try:
        phdf.lib()
    except RuntimeError as e:
        pytest.skip(str(e))
    return phdf


def read_dataset(path, name):
    r = subprocess.run([H5DUMP, "-d", name, "-m", "%.17g", "-w", "0", "-y", path],
                       capture_output=True, text=True, check=True)
    dims = re.search(r"DATASPACE\s+SIMPLE\s+\{\s*\(([^)]*)\)", r.stdout)
    shape = tuple(int(x) for x in dims.group(1).split(","))
    body = r.stdout[r.stdout.index("DATA {") + 6:]
    body = body[:body.index("}")]
    vals = np.array([float(x) for x in body.replace("\n", " ").split(",") if x.strip()])
    return vals.reshape(shape)


def read_attr(path, obj, name):
    r = subprocess.run([H5DUMP, "-a", f"{obj}/{name}", "-m", "%.17g", "-w", "0", "-y", path],
                       capture_output=True, text=True, check=True)
    body = r.stdout[r.stdout.index("DATA {") + 6:]
    body = body[:body.index("}")]
    return [x.strip().strip('"') for x in body.replace("\n", " ").split(",") if x.strip()]


def xzplot_assemble(path, field):
    """xzplot.py:33-80 restated: block maps averaged over x3 and placed by
    LogicalLocations, coordinates from VolumeLocations."""
    data = read_dataset(path, "/" + field)
    rho_blocks = data.mean(axis=1)
    x1b = read_dataset(path, "/VolumeLocations/x")
    x2b = read_dataset(path, "/VolumeLocations/y")
    ll = read_dataset(path, "/LogicalLocations").astype(int)
    nb, nyb, nxb = rho_blocks.shape
    nx, ny = (ll[:, 0].max() + 1) * nxb, (ll[:, 1].max() + 1) * nyb
    q = np.zeros((ny, nx))
    x1, x2 = np.zeros(nx), np.zeros(ny)
    for b in range(nb):
        i0, j0 = ll[b, 0] * nxb, ll[b, 1] * nyb
        q[j0:j0 + nyb, i0:i0 + nxb] = rho_blocks[b]
        x1[i0:i0 + nxb] = x1b[b]
        x2[j0:j0 + nyb] = x2b[b]
    ix, iy = np.argsort(x1), np.argsort(x2)
    return x1[ix], x2[iy], q[np.ix_(iy, ix)]


def _synthetic(tmp_path, phdf):
    """2 x 2 blocks of 6 x 4 x 3 cells, a known global field."""
    nbx, nby, nxb = (2, 2), None, (6, 4, 3)
    root_min, dxs = np.array([0.5, -1.0, 0.0]), np.array([0.25, 0.5, 2.0])
    lx = np.array([[i, j, 0] for j in range(2) for i in range(2)], dtype=np.int64)
    nb = len(lx)
    xmin = np.array([root_min + lx[b] * np.array(nxb) * dxs for b in range(nb)])
    dx = np.tile(dxs, (nb, 1))
    # global field f(i, j, k) = i + 100 j + 10000 k
    dens = np.zeros((nb, nxb[2], nxb[1], nxb[0]))
    vel = np.zeros((nb, 3, nxb[2], nxb[1], nxb[0]))
    for b in range(nb):
        for k in range(nxb[2]):
            for j in range(nxb[1]):
                for i in range(nxb[0]):
                    gi, gj = lx[b, 0] * nxb[0] + i, lx[b, 1] * nxb[1] + j
                    dens[b, k, j, i] = gi + 100 * gj + 10000 * k
                    vel[b, :, k, j, i] = (gi * 0.5, gj * 0.25, -k)
    m = phdf.Mesh()
    m.nblocks = nb
    for d in range(3):
        m.nx[d] = nxb[d]
        m.root_nx[d] = nxb[d] * (2 if d < 2 else 1)
        m.root_min[d] = root_min[d]
        m.root_max[d] = root_min[d] + m.root_nx[d] * dxs[d]
    m.ng, m.ndim, m.ncycle, m.time, m.dt = 4, 3, 7, 1.25, 0.01
    m.lx = lx.ctypes.data_as(C.POINTER(C.c_int64))
    m.xmin = xmin.ctypes.data_as(C.POINTER(C.c_double))
    m.dx = dx.ctypes.data_as(C.POINTER(C.c_double))
    vs = (phdf.Var * 2)()
    vs[0] = phdf.Var(b"density", 1, dens.ctypes.data_as(C.POINTER(C.c_double)))
    vs[1] = phdf.Var(b"weighted_velocity", 3, vel.ctypes.data_as(C.POINTER(C.c_double)))
    path = str(tmp_path / "synthetic.phdf")
    assert phdf.lib().pangu_write_phdf(path.encode(), C.byref(m), 2, vs) == 0
    return path, dens, vel, lx, xmin, dxs, nxb


def test_layout_and_values_round_trip(tmp_path):
    phdf = _phdf()
    path, dens, vel, lx, xmin, dxs, nxb = _synthetic(tmp_path, phdf)
    np.testing.assert_array_equal(read_dataset(path, "/density"), dens)
    np.testing.assert_array_equal(read_dataset(path, "/weighted_velocity"), vel)
    np.testing.assert_array_equal(read_dataset(path, "/LogicalLocations"), lx)
    np.testing.assert_array_equal(read_dataset(path, "/Blocks/loc.lx123"), lx)
    np.testing.assert_array_equal(read_dataset(path, "/Levels"), np.zeros(len(lx)))
    for d, ax in enumerate("xyz"):
        cen = read_dataset(path, "/VolumeLocations/" + ax)
        face = read_dataset(path, "/Locations/" + ax)
        for b in range(len(lx)):
            np.testing.assert_array_equal(cen[b], xmin[b, d] + (np.arange(nxb[d]) + 0.5) * dxs[d])
            np.testing.assert_array_equal(face[b], xmin[b, d] + np.arange(nxb[d] + 1) * dxs[d])
    assert float(read_attr(path, "/Info", "Time")[0]) == 1.25
    assert int(read_attr(path, "/Info", "NCycle")[0]) == 7
    assert read_attr(path, "/Info", "OutputDatasetNames") == ["density", "weighted_velocity"]
    assert [int(x) for x in read_attr(path, "/Info", "MeshBlockSize")] == list(nxb)


def test_levels_of_a_refined_mesh(tmp_path):
    """Per-block refinement levels go to /Levels and their maximum to
    /Info MaxLevel (NULL levels write zeros, as the round trip above)."""
    phdf = _phdf()
    nb, nxb = 3, (4, 4, 1)
    lx = np.array([[0, 0, 0], [2, 0, 0], [3, 1, 0]], dtype=np.int64)
    level = np.array([0, 1, 1], dtype=np.int32)
    xmin = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [1.5, 0.5, 0.0]])
    dx = np.array([[0.25, 0.25, 1.0], [0.125, 0.125, 1.0], [0.125, 0.125, 1.0]])
    dens = np.arange(nb * 16, dtype=np.float64).reshape(nb, 1, 4, 4)
    m = phdf.Mesh()
    m.nblocks = nb
    for d in range(3):
        m.nx[d] = nxb[d]
        m.root_nx[d] = (8, 4, 1)[d]
        m.root_max[d] = (2.0, 1.0, 1.0)[d]
    m.ng, m.ndim = 4, 2
    m.lx = lx.ctypes.data_as(C.POINTER(C.c_int64))
    m.xmin = xmin.ctypes.data_as(C.POINTER(C.c_double))
    m.dx = dx.ctypes.data_as(C.POINTER(C.c_double))
    m.level = level.ctypes.data_as(C.POINTER(C.c_int))
    vs = (phdf.Var * 1)()
    vs[0] = phdf.Var(b"density", 1, dens.ctypes.data_as(C.POINTER(C.c_double)))
    path = str(tmp_path / "refined.phdf")
    assert phdf.lib().pangu_write_phdf(path.encode(), C.byref(m), 1, vs) == 0
    np.testing.assert_array_equal(read_dataset(path, "/Levels"), level)
    np.testing.assert_array_equal(read_dataset(path, "/LogicalLocations"), lx)
    assert int(read_attr(path, "/Info", "MaxLevel")[0]) == 1
    np.testing.assert_array_equal(read_dataset(path, "/VolumeLocations/x")[2], 1.5 + (np.arange(4) + 0.5) * 0.125)


def test_reference_plot_script_assembles_the_global_field(tmp_path):
    phdf = _phdf()
    path, dens, vel, lx, xmin, dxs, nxb = _synthetic(tmp_path, phdf)
    x1, x2, q = xzplot_assemble(path, "density")
    gi, gj = np.meshgrid(np.arange(12), np.arange(8), indexing="xy")
    want = gi + 100 * gj + 10000 * np.mean(np.arange(nxb[2]))
    np.testing.assert_allclose(q, want, rtol=0, atol=1e-9)
    np.testing.assert_allclose(x1, 0.5 + (np.arange(12) + 0.5) * 0.25)
    np.testing.assert_allclose(x2, -1.0 + (np.arange(8) + 0.5) * 0.5)


def test_xdmf_references_existing_datasets(tmp_path):
    phdf = _phdf()
    path, *_ = _synthetic(tmp_path, phdf)
    root = ET.parse(path + ".xdmf").getroot()
    grids = root.findall(".//Grid[@GridType='Uniform']")
    assert len(grids) == 4
    refs = {e.text.split(":")[1] for e in root.iter("DataItem") if e.get("Format") == "HDF"}
    assert refs == {"/Locations/x", "/Locations/y", "/Locations/z", "/density", "/weighted_velocity"}
    names = [a.get("Name") for a in grids[0].findall("Attribute")]
    assert names == ["density", "weighted_velocity_0", "weighted_velocity_1", "weighted_velocity_2"]
    r = subprocess.run([H5DUMP, "-H", path], capture_output=True, text=True, check=True)
    for ref in refs:
        assert f'DATASET "{ref.split("/")[-1]}"' in r.stdout


def test_writer_rejects_bad_input(tmp_path):
    phdf = _phdf()
    m = phdf.Mesh()
    assert phdf.lib().pangu_write_phdf(str(tmp_path / "x.phdf").encode(), C.byref(m), 0, None) == -1
    assert b"no blocks" in phdf.lib().pangu_phdf_last_error()


@pytest.mark.gpu
def test_torus_output_matches_the_gpu_fields(tmp_path):
    phdf = _phdf()
    from pangu_amd import configs
    from pangu_amd.sim import HipSim
    d = configs.torus2d(64, 32, (2, 2, 1))
    h = HipSim(d)
    h.init()
    h.step(2, "graph")
    path = phdf.write(h, str(tmp_path / "torus.phdf"), ncycle=2)
    P = h.interior(h.get("prim", 0))
    np.testing.assert_array_equal(read_dataset(path, "/density"), P[:, 0])
    np.testing.assert_array_equal(read_dataset(path, "/magnetic_field"), P[:, 5:8])
    np.testing.assert_array_equal(read_dataset(path, "/electron_entropy"), P[:, 9])
    x1, x2, q = xzplot_assemble(path, "density")
    assert q.shape == (32, 64)
    np.testing.assert_allclose(np.diff(x1), (6.0 - 0.2) / 64, rtol=1e-12)
    assert float(read_attr(path, "/Info", "Time")[0]) == h.mesh_state(0)[1]
    h.close()


def test_library_exports_every_declared_function():
    phdf = _phdf()
    hdr = open(os.path.join(os.path.dirname(__file__), "..", "include", "pangu_phdf.h")).read()
    names = re.findall(r"^\w[\w\s\*]*?\b(pangu_\w+)\s*\(", hdr, re.M)
    assert set(names) == {"pangu_write_phdf", "pangu_phdf_last_error"}
    for n in names:
        assert hasattr(phdf.lib(), n)
