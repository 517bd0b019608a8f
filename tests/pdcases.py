"""Inputs of the reference's own test models, rebuilt from golden data.

The 10-step regression model of peripy/test/conftest.py:66-84 /
test_regression.py:11-23: example_mesh.vtk (2113 nodes), horizon 0.1,
s0 = 0.005, k = 18*0.05/(pi*0.1**4), Euler dt = 1e-3, displacement BC
magnitudes 1e-5/2 * linspace(1, 10, 10), the pre-cracked connectivity of
expected_connectivity_crack(_cl).npz, family from
expected_connectivity_basic.npz, volume from expected_volume.npy.
"""
import pathlib

import numpy as np

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"

HORIZON = 0.1
CRITICAL_STRETCH = 0.005
BOND_STIFFNESS = 18.0 * 0.05 / (np.pi * 0.1 ** 4)
DT = 1e-3
STEPS = 10
DISP_MAGS = 0.00001 / 2 * np.linspace(1, STEPS, STEPS)


def read_vtk_points(path):
    """POINTS of a meshio-written BINARY legacy VTK (big-endian doubles)."""
    raw = pathlib.Path(path).read_bytes()
    tag = b"POINTS "
    p = raw.index(tag)
    e = raw.index(b"\n", p)
    n = int(raw[p + len(tag):e].split()[0])
    return np.frombuffer(raw, dtype=">f8", count=3 * n,
                         offset=e + 1).reshape(n, 3).astype(np.float64)


def simple_displacement_boundary(x):
    """peripy/test/conftest.py:24-40."""
    bnd = [None, None, None]
    if x[0] < 1.5 * 0.1:
        bnd[0] = -1
        bnd[1] = 0
        bnd[2] = 0
    elif x[0] > 1.0 - 1.5 * 0.1:
        bnd[0] = 1
        bnd[1] = 0
        bnd[2] = 0
    return bnd


def is_crack(x, y):
    """peripy/test/conftest.py:43-63 (the predicate, before the helper)."""
    crack_length = 0.3
    output = 0
    p1 = x
    p2 = y
    if x[0] > y[0]:
        p2 = x
        p1 = y
    if p1[0] < 0.5 + 1e-6 and p2[0] > 0.5 + 1e-6:
        m = (p2[1] - p1[1]) / (p2[0] - p1[0])
        c = p1[1] - m * p1[0]
        height = m * 0.5 + c
        if (height > 0.5 * (1 - crack_length)
                and height < 0.5 * (1 + crack_length)):
            output = 1
    return output


def regression_inputs(layout="cython"):
    """Arrays of the 10-step regression model.

    layout "cython": nlist (2113, 68) compacted, 0-padded;
    layout "cl": nlist (2113, 128), -1 marked.
    """
    coords = read_vtk_points(GOLDEN / "example_mesh.vtk")
    n = coords.shape[0]
    volume = np.load(GOLDEN / "expected_volume.npy")
    family = np.load(GOLDEN / "expected_connectivity_basic.npz")["n_neigh"]
    name = ("expected_connectivity_crack.npz" if layout == "cython"
            else "expected_connectivity_crack_cl.npz")
    z = np.load(GOLDEN / name)
    nlist = np.ascontiguousarray(z["nlist"], dtype=np.int32)
    n_neigh = np.ascontiguousarray(z["n_neigh"], dtype=np.int32)
    bc_types = np.zeros((n, 3), dtype=np.int32)
    bc_values = np.zeros((n, 3), dtype=np.float64)
    for i in range(n):
        b = simple_displacement_boundary(coords[i])
        for d in range(3):
            if b[d] is not None:
                bc_types[i, d] = 1
                bc_values[i, d] = b[d]
    return dict(coords=coords, volume=volume.astype(np.float64),
                family=family.astype(np.int32), nlist=nlist,
                n_neigh=n_neigh, bc_types=bc_types, bc_values=bc_values,
                force_bc_types=np.zeros((n, 3), dtype=np.int32),
                force_bc_values=np.zeros((n, 3), dtype=np.float64))


def rel_l2(a, b):
    """||a - b|| / ||b|| (0 when both vanish)."""
    nb = np.linalg.norm(b)
    na = np.linalg.norm(np.asarray(a) - np.asarray(b))
    return na / nb if nb > 0 else na


def lattice(shape, dx, jitter=0.0, seed=None):
    """Cell-centred lattice points, x slowest, z fastest; optional
    seeded uniform jitter of +-jitter*dx per coordinate."""
    nx, ny, nz = shape
    g = np.stack(np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz),
                             indexing="ij"), -1).reshape(-1, 3)
    x = (g + 0.5) * dx
    if jitter:
        rng = np.random.default_rng(seed)
        x = x + rng.uniform(-jitter * dx, jitter * dx, size=x.shape)
    return np.ascontiguousarray(x, dtype=np.float64)
