"""GPU neighbour list (pd_neighbour_list_device) against the host search."""
import time

import numpy as np
import pytest

import pdcases as pc
from probabilistic_peridynamics_amd import _lib, configs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    try:
        ok = _lib.device_count() > 0
    except OSError:
        ok = False
    if not ok:
        pytest.fail("no HIP device: the gpu tests need an MI355X")


def _both(coords, horizon, k=None, pad=-1):
    fh, nh = _lib.neighbour_list(coords, horizon, k, pad)
    fd, nd = _lib.neighbour_list(coords, horizon, k, pad, device=0)
    np.testing.assert_array_equal(fd, fh)
    np.testing.assert_array_equal(nd, nh)
    return fh, nh


def test_example_mesh(golden):
    coords = pc.read_vtk_points(golden / "example_mesh.vtk")
    _both(coords, 0.1, pad=0)


@pytest.mark.parametrize("seed,shape", [(0, (30, 12, 9)), (1, (7, 41, 5))])
def test_jittered_lattices(seed, shape):
    dx = 0.01
    coords = configs.lattice(shape, dx, jitter=0.3, seed=seed)
    fam, nl = _both(coords, 3.01 * dx)
    assert fam.min() > 0
    _both(coords, 2.2 * dx, k=int(fam.max()) + 3, pad=0)


def test_family_only_and_overflow():
    coords = configs.lattice((12, 6, 6), 0.01, jitter=0.2, seed=4)
    fh = _lib.family_count(coords, 0.0301)
    fd = _lib.family_count(coords, 0.0301, device=0)
    np.testing.assert_array_equal(fd, fh)
    with pytest.raises(_lib.PDError):
        _lib.neighbour_list(coords, 0.0301, int(fh.max()) - 1, -1, device=0)


def test_full_size_beam():
    w = configs.beam_vv()
    t0 = time.perf_counter()
    fd, nd = _lib.neighbour_list(w.coords, w.horizon, w.max_neighbours, -1,
                                 device=0)
    t_dev = time.perf_counter() - t0
    print(f"device neighbour list: {t_dev:.2f} s for {w.nnodes} nodes")
    np.testing.assert_array_equal(fd, w.family)
    np.testing.assert_array_equal(nd, w.nlist)
