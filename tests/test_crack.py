"""Initial crack (create_crack.pyx, initial_crack_helper model.py:1578-1612):
the library's create_crack against the oracle, and the bulk LineCrack
predicate (host path) against the reference's pairwise test predicate and
its post-crack fixtures (CPU only)."""
import warnings

import numpy as np
import pytest

import oracle
import pdcases as pc
from probabilistic_peridynamics_amd.crack import LineCrack, create_crack
from probabilistic_peridynamics_amd.model import Model, initial_crack_helper

from test_host import Recorder

# conftest.py:43-63 as a LineCrack: crack length 0.3 across x = 0.5
REF_CRACK = dict(axis=0, cross_axis=1, split=0.5 + 1e-6, at=0.5,
                 lo=0.5 * (1 - 0.3), hi=0.5 * (1 + 0.3))


def _basic(golden, layout):
    coords = pc.read_vtk_points(golden / "example_mesh.vtk")
    basic = np.load(golden / "expected_connectivity_basic.npz")
    fam = basic["n_neigh"].astype(np.int32)
    if layout == "cython":
        nl = basic["nlist"].astype(np.int32).copy()
    else:
        nl = -np.ones((len(fam), 128), dtype=np.int32)
        for i in range(len(fam)):
            nl[i, :fam[i]] = basic["nlist"][i, :fam[i]]
    return coords, nl, fam.copy()


def test_predicate_equals_reference_predicate(golden):
    coords, nl, nn = _basic(golden, "cython")
    lc = LineCrack(**REF_CRACK)
    hits = 0
    for i in range(len(nn)):
        for j in nl[i, :nn[i]]:
            a = lc.predicate(coords[i], coords[j])
            assert a == pc.is_crack(coords[i], coords[j])
            hits += a
    # every cracked bond is seen from both ends, and the crack removes
    # exactly those bonds from the reference's lists
    z = np.load(golden / "expected_connectivity_crack.npz")
    assert hits > 0 and hits == int(nn.sum()) - int(z["n_neigh"].sum())


@pytest.mark.parametrize("layout,name", [
    ("cython", "expected_connectivity_crack.npz"),
    ("cl", "expected_connectivity_crack_cl.npz")])
def test_line_crack_gives_the_crack_fixtures(golden, layout, name):
    """Bulk pairs (host path) = initial_crack_helper(is_crack) pairs, in
    order; create_crack of them = the reference's post-crack lists."""
    coords, nl, nn = _basic(golden, layout)
    pairs = LineCrack(**REF_CRACK, device=-1)(coords, nl, nn)
    ref = initial_crack_helper(pc.is_crack)(coords, nl, nn)
    assert np.array_equal(pairs, np.array(ref, dtype=np.int32).reshape(-1, 2))
    create_crack(pairs, nl, nn)
    z = np.load(golden / name)
    assert np.array_equal(nl, z["nlist"])
    assert np.array_equal(nn, z["n_neigh"])


def test_model_with_line_crack(golden):
    """Model(initial_crack=LineCrack) = Model(initial_crack=
    initial_crack_helper(is_crack)) bitwise, and the reference's post-crack
    bond sets (the model's rows are sorted, the fixture's are in the
    reference's KD-tree order)."""
    def model(crack):
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")
            return Model(golden / "example_mesh.vtk", Recorder(1e-3, "cl"),
                         horizon=0.1, critical_stretch=0.005,
                         bond_stiffness=1.0, initial_crack=crack)
    nl, nn = model(LineCrack(**REF_CRACK, device=-1)).initial_connectivity
    nl2, nn2 = model(initial_crack_helper(pc.is_crack)).initial_connectivity
    assert np.array_equal(nl, nl2) and np.array_equal(nn, nn2)
    z = np.load(golden / "expected_connectivity_crack_cl.npz")
    assert np.array_equal(nn, z["n_neigh"])
    for i in range(len(nn)):
        assert set(nl[i, :nn[i]]) == set(z["nlist"][i, :nn[i]])


def test_create_crack_equals_oracle_on_random_lists():
    """Repeated pairs, pairs that are not bonds, repeats inside rows."""
    rng = np.random.default_rng(5)
    for trial in range(200):
        n = int(rng.integers(1, 30))
        k = int(rng.integers(1, 10))
        nl = rng.integers(-1, n, size=(n, k)).astype(np.int32)
        nn = rng.integers(0, k + 1, size=n).astype(np.int32)
        crack = rng.integers(0, n, size=(int(rng.integers(0, 40)), 2))
        a, an = nl.copy(), nn.copy()
        b, bn = nl.copy(), nn.copy()
        oracle.create_crack(crack, a, an)
        create_crack(crack, b, bn)
        assert np.array_equal(a, b) and np.array_equal(an, bn), trial


def test_create_crack_rejects_bad_pairs_before_changing_anything():
    nl = np.array([[1, -1], [0, -1]], dtype=np.int32)
    nn = np.array([1, 1], dtype=np.int32)
    with pytest.raises(ValueError, match="outside"):
        create_crack([(0, 1), (0, 2)], nl, nn)
    assert np.array_equal(nl, [[1, -1], [0, -1]])
    with pytest.raises(ValueError, match="int32"):
        create_crack([(0, 1)], nl.astype(np.int64), nn)


def test_line_crack_other_axes_and_capacity():
    """A crack across z = 0.5 in the (z, x) plane of a jittered lattice: the
    bulk pairs equal the pairwise predicate run by initial_crack_helper."""
    from probabilistic_peridynamics_amd import _lib, configs
    coords = configs.lattice((9, 7, 10), 0.1, jitter=0.2, seed=4)
    fam, nl = _lib.neighbour_list(coords, 0.25, None, -1)
    lc = LineCrack(2, 0, split=0.5, at=0.5, lo=0.2, hi=0.6, device=-1)
    pairs = lc(coords, nl, fam)
    ref = initial_crack_helper(lc.predicate)(coords, nl, fam)
    assert len(ref) > 0
    assert np.array_equal(pairs, np.array(ref, dtype=np.int32).reshape(-1, 2))
    with pytest.raises(ValueError, match="two of"):
        LineCrack(1, 1, 0, 0, 0, 0)
