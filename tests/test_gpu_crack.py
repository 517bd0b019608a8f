"""Initial crack on the GPU (pd_line_crack_pairs, device 0) against the host
path, the reference's pairwise predicate and a numpy restatement of it."""
import time

import numpy as np
import pytest

import pdcases as pc
from probabilistic_peridynamics_amd import _lib, configs
from probabilistic_peridynamics_amd.crack import LineCrack, create_crack
from probabilistic_peridynamics_amd.model import initial_crack_helper

pytestmark = pytest.mark.gpu

REF_CRACK = dict(axis=0, cross_axis=1, split=0.5 + 1e-6, at=0.5,
                 lo=0.5 * (1 - 0.3), hi=0.5 * (1 + 0.3))


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    try:
        ok = _lib.device_count() > 0
    except OSError:
        ok = False
    if not ok:
        pytest.fail("no HIP device: the gpu tests need an MI355X")


def _pairs_numpy(lc, coords, nlist, n_neigh):
    """The predicate over every bond i < j at once, in helper order."""
    n, k = nlist.shape
    slot = np.arange(k)[None, :] < n_neigh[:, None]
    i = np.broadcast_to(np.arange(n)[:, None], (n, k))[slot]
    j = nlist[slot]
    keep = j > i
    i, j = i[keep], j[keep]
    a, c = lc.axis, lc.cross_axis
    xi, xj = coords[i], coords[j]
    swap = xi[:, a] > xj[:, a]
    p1 = np.where(swap[:, None], xj, xi)
    p2 = np.where(swap[:, None], xi, xj)
    cross = (p1[:, a] < lc.split) & (p2[:, a] > lc.split)
    p1, p2, i, j = p1[cross], p2[cross], i[cross], j[cross]
    m = (p2[:, c] - p1[:, c]) / (p2[:, a] - p1[:, a])
    b = p1[:, c] - m * p1[:, a]
    h = m * lc.at + b
    hit = (h > lc.lo) & (h < lc.hi)
    return np.stack([i[hit], j[hit]], axis=1).astype(np.int32)


def test_example_mesh(golden):
    coords = pc.read_vtk_points(golden / "example_mesh.vtk")
    fam, nl = _lib.neighbour_list(coords, 0.1, 128, -1)
    lc = LineCrack(**REF_CRACK, device=0)
    pairs = lc(coords, nl, fam)
    ref = initial_crack_helper(pc.is_crack)(coords, nl, fam)
    assert np.array_equal(pairs, np.array(ref, dtype=np.int32).reshape(-1, 2))


@pytest.mark.parametrize("axis,cross", [(0, 1), (1, 2), (2, 0)])
def test_jittered_lattice_all_axes(axis, cross):
    coords = configs.lattice((21, 17, 13), 0.05, jitter=0.25, seed=axis)
    fam, nl = _lib.neighbour_list(coords, 0.151, None, -1)
    lc = LineCrack(axis, cross, split=0.5, at=0.5, lo=0.1, hi=0.45)
    dev = LineCrack(axis, cross, 0.5, 0.5, 0.1, 0.45, device=0)(
        coords, nl, fam)
    host = LineCrack(axis, cross, 0.5, 0.5, 0.1, 0.45, device=-1)(
        coords, nl, fam)
    assert len(dev) > 0
    assert np.array_equal(dev, host)
    assert np.array_equal(dev, _pairs_numpy(lc, coords, nl, fam))


def test_full_size_beam_notch():
    """A notch across the 1,038,230-node beam at mid-span, from the bottom
    face to half depth."""
    w = configs.beam_vv()
    nl = np.ascontiguousarray(w.nlist, dtype=np.int32)
    nn = np.ascontiguousarray(w.n_neigh, dtype=np.int32)
    args = (0, 2, 1.0 + 1e-9, 1.0, -1.0, 0.1)
    t0 = time.perf_counter()
    dev = LineCrack(*args, device=0)(w.coords, nl, nn)
    t_dev = time.perf_counter() - t0
    t0 = time.perf_counter()
    host = LineCrack(*args, device=-1)(w.coords, nl, nn)
    t_host = time.perf_counter() - t0
    print(f"line crack pairs: GPU {t_dev:.3f} s, host {t_host:.3f} s, "
          f"{len(dev)} pairs")
    assert len(dev) > 1000
    assert np.array_equal(dev, host)
    assert np.array_equal(dev, _pairs_numpy(LineCrack(*args), w.coords, nl,
                                            nn))
    before = int(nn.sum())
    create_crack(dev, nl, nn)
    assert int(nn.sum()) == before - 2 * len(dev)
