"""State I/O on the host: the VTK legacy writer (efvm_write_vtk, the reference's
crates/cfd_output/src/lib.rs:36-113) checked for structure as the reference's own test
writes_well_formed_vtk does (:168-200), and its values against the oracle's EOS."""
import re

import numpy as np
import pytest

import dev_field_efvm_amd as efvm
from oracle import orc


def parse_vtk(path):
    text = open(path).read()
    blocks = {}
    name = None
    for line in text.splitlines():
        m = re.match(r"SCALARS (\w+) double 1", line)
        if m:
            name = m.group(1)
            blocks[name] = []
            continue
        if line.startswith("VECTORS velocity"):
            name = "velocity"
            blocks[name] = []
            continue
        if line.startswith("LOOKUP_TABLE"):
            continue
        if name and re.match(r"^-?\d", line):
            blocks[name].append([float(x) for x in line.split()])
    return text, {k: np.array(v) for k, v in blocks.items()}


def test_vtk_is_well_formed(tmp_path):
    nx, ny = 4, 3
    u = orc.prim_to_cons(np.array([1.0, 10.0, 0.0, 0.0, 1e5, 348.0]), 1.4)
    st = np.tile(u, (ny, nx, 1))
    path = str(tmp_path / "out.vtk")
    efvm.write_vtk(path, st, nx, ny, hi=(1.0, 1.0), dz=1.0)
    text, b = parse_vtk(path)
    assert text.startswith("# vtk DataFile Version 3.0\n")
    assert "DATASET RECTILINEAR_GRID" in text
    assert "DIMENSIONS 5 4 2" in text                 # (ni+1, nj+1, nk+1), one cell thick
    assert "CELL_DATA 12" in text
    assert "X_COORDINATES 5 double" in text and "Y_COORDINATES 4 double" in text and "Z_COORDINATES 2 double" in text
    for k in ("density", "pressure", "temperature", "mach"):
        assert b[k].shape == (12, 1), k
    assert b["velocity"].shape == (12, 3)
    assert np.allclose(b["pressure"], 1e5, rtol=1e-9)
    xs = re.search(r"X_COORDINATES 5 double\n((?:.*\n){5})", text).group(1).split()
    assert np.allclose([float(x) for x in xs], [0.0, 0.25, 0.5, 0.75, 1.0], atol=1e-15)


def test_vtk_values_follow_the_eos(tmp_path):
    rng = np.random.default_rng(3)
    nx, ny, gamma, r_gas = 5, 4, 1.4, 287.058
    st = np.zeros((ny, nx, 5))
    for j in range(ny):
        for i in range(nx):
            P = np.array([rng.uniform(0.5, 2.0), rng.uniform(-50, 50), rng.uniform(-50, 50), 0.0,
                          rng.uniform(0.5e5, 2e5), 300.0])
            st[j, i] = orc.prim_to_cons(P, gamma)
    st[1, 2] = [0.0, 0.0, 0.0, 0.0, 0.0]              # rho = 0: velocity and Mach are not finite
    path = str(tmp_path / "v.vtk")
    efvm.write_vtk(path, st, nx, ny, lo=(-1.0, 2.0), hi=(4.0, 6.0), dz=0.5, gamma=gamma, r_gas=r_gas)
    text, b = parse_vtk(path)
    k = 0
    for j in range(ny):
        for i in range(nx):
            u = st[j, i]
            if u[0] == 0.0:
                assert b["velocity"][k].tolist() == [0.0, 0.0, 0.0]     # non-finite written as zero
                assert b["mach"][k, 0] == 0.0
            else:
                P = orc.cons_to_prim(u, gamma, r_gas)
                assert np.isclose(b["density"][k, 0], u[0], rtol=1e-10)
                assert np.isclose(b["pressure"][k, 0], P[4], rtol=1e-9)
                assert np.isclose(b["temperature"][k, 0], P[5], rtol=1e-9)
                assert np.allclose(b["velocity"][k], P[1:4], rtol=1e-9, atol=1e-12)
                mach = np.linalg.norm(P[1:4]) / orc.sound_speed(u, gamma)
                assert np.isclose(b["mach"][k, 0], mach, rtol=1e-9)
            k += 1
    # signed, at least two-digit exponents, as the reference's fmt_e writes them
    assert re.fullmatch(r"-?\d\.\d{10}e[+-]\d{2}", text.splitlines()[6]), text.splitlines()[6]


def test_vtk_3d_grid_and_cell_order(tmp_path):
    """efvm3_write_vtk: the same writer over an nx x ny x nz mesh; cells x fastest, then y, then z
    (the VTK rectilinear-grid cell order and the wire order), with rho_w in the velocity."""
    nx, ny, nz, gamma = 3, 2, 4, 1.4
    st = np.zeros((nz, ny, nx, 5))
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                P = np.array([1.0 + 0.1 * i + 0.01 * j + 0.001 * k, 5.0 * i, -3.0 * j, 2.0 * k, 1e5, 300.0])
                st[k, j, i] = orc.prim_to_cons(P, gamma)
    path = str(tmp_path / "v3.vtk")
    efvm.write_vtk3(path, st, nx, ny, nz, lo=(0.0, -1.0, 2.0), hi=(3.0, 1.0, 6.0), gamma=gamma)
    text, b = parse_vtk(path)
    assert "DIMENSIONS 4 3 5" in text and "CELL_DATA 24" in text
    zs = re.search(r"Z_COORDINATES 5 double\n((?:.*\n){5})", text).group(1).split()
    assert np.allclose([float(z) for z in zs], [2.0, 3.0, 4.0, 5.0, 6.0], atol=1e-14)
    q = 0
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                assert np.isclose(b["density"][q, 0], 1.0 + 0.1 * i + 0.01 * j + 0.001 * k, rtol=1e-10)
                assert np.allclose(b["velocity"][q], [5.0 * i, -3.0 * j, 2.0 * k], rtol=1e-9, atol=1e-12)
                q += 1


def test_vtk_3d_stretched_axes(tmp_path):
    """efvm3_write_vtk_edges: a stretched axis's coordinates are each cell's centroid minus half
    its spacing, then the last centroid plus half its spacing (crates/cfd_output/src/lib.rs:50-72),
    which is the edge list up to rounding; bad edges are refused."""
    nx, ny, nz = 2, 3, 4
    st = np.broadcast_to(orc.prim_to_cons([1.0, 1.0, 0, 0, 1e5, 300.0], 1.4), (nz, ny, nx, 5)).copy()
    ey, ez = [-1.0, -0.7, 0.2, 1.0], [2.0, 2.1, 2.5, 4.0, 6.0]
    path = str(tmp_path / "s3.vtk")
    efvm.write_vtk3(path, st, nx, ny, nz, lo=(0.0, -1.0, 2.0), hi=(3.0, 1.0, 6.0), y_edges=ey, z_edges=ez)
    text, _ = parse_vtk(path)
    ys = [float(v) for v in re.search(r"Y_COORDINATES 4 double\n((?:.*\n){4})", text).group(1).split()]
    zs = [float(v) for v in re.search(r"Z_COORDINATES 5 double\n((?:.*\n){5})", text).group(1).split()]
    xs = [float(v) for v in re.search(r"X_COORDINATES 3 double\n((?:.*\n){3})", text).group(1).split()]
    assert np.allclose(ys, ey, rtol=0, atol=1e-15) and np.allclose(zs, ez, rtol=0, atol=1e-15)
    assert np.allclose(xs, [0.0, 1.5, 3.0], rtol=0, atol=1e-15)
    with pytest.raises(efvm.EfvmError):
        efvm.write_vtk3(path, st, nx, ny, nz, lo=(0.0, -1.0, 2.0), hi=(3.0, 1.0, 6.0), y_edges=[-1.0, 0.5, 0.2, 1.0])
    with pytest.raises(efvm.EfvmError):
        efvm.write_vtk3(path, st, nx, ny, nz, lo=(0.0, -1.0, 2.0), hi=(3.0, 1.0, 6.0), z_edges=[2.0, 2.1, 2.5, 4.0, 5.0])
