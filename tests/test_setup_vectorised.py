"""Set-up callbacks evaluated for all bonds / nodes at once
(:func:`vectorised`) against the reference's per-bond and per-node calls
(model.py:657-780) and its bond-type fixture (test_model.py:543-561)."""
import warnings

import numpy as np
import pytest

from probabilistic_peridynamics_amd.model import Model, vectorised

from test_host import Recorder


def bond_type(x, y):
    """test_model.py:548-552."""
    if x[0] == 0.0:
        return 0
    else:
        return 1


@vectorised
def bond_type_v(x, y):
    return np.where(x[0] == 0.0, 0, 1)


def _model(golden, layout="cl", **kw):
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        return Model(golden / "example_mesh.vtk", Recorder(1e-3, layout),
                     horizon=0.1, critical_stretch=[[0.05], [0.05]],
                     bond_stiffness=[[1.0], [2.0]], **kw)


@pytest.fixture(scope="module")
def expected(golden):
    return dict(np.load(golden / "expected_bond_types.npz"))


@pytest.mark.parametrize("fn", [bond_type, bond_type_v])
def test_bond_types_match_fixture(golden, expected, fn):
    """This bond type depends on node i only, so the fixture holds slot by
    slot whatever the row order."""
    m = _model(golden, is_bond_type=fn)
    assert np.array_equal(m.bond_types, expected["expected_bond_types"])
    assert np.array_equal(m.bond_types, expected["expected_bond_types_cl"])


def test_vectorised_equals_per_bond_calls_in_blocks(golden):
    """A bond type that depends on both ends, evaluated in blocks smaller
    than one row and larger than many."""
    def f(x, y):
        return int(x[0] + 0.3 * y[1] > 0.6) + int(y[2] > 10.0)

    @vectorised
    def fv(x, y):
        return ((x[0] + 0.3 * y[1] > 0.6).astype(np.intc)
                + (y[2] > 10.0).astype(np.intc))

    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        m = Model(golden / "example_mesh.vtk", Recorder(1e-3),
                  horizon=0.1, critical_stretch=[[0.05], [0.05], [0.05]],
                  bond_stiffness=[[1.0], [2.0], [3.0]], is_bond_type=f)
    ref = m.bond_types
    nl, nn = m.initial_connectivity
    for block in (7, 1000, 1 << 22):
        out = np.zeros_like(ref)
        m._bond_types_vectorised(fv, nl, nn, 3, out, block=block)
        assert np.array_equal(out, ref), block


def test_vectorised_bond_type_errors(golden):
    with pytest.raises(ValueError, match="positive"):
        _model(golden, is_bond_type=vectorised(
            lambda x, y: np.full(x.shape[1], -1)))
    with pytest.raises(ValueError, match="less"):
        _model(golden, is_bond_type=vectorised(
            lambda x, y: np.full(x.shape[1], 2)))
    with pytest.raises(TypeError, match="integer array"):
        _model(golden, is_bond_type=vectorised(
            lambda x, y: np.zeros(x.shape[1])))
    with pytest.raises(TypeError, match="integer array"):
        _model(golden, is_bond_type=vectorised(lambda x, y: 0))


def test_vectorised_density(golden):
    def rho(x):
        return float(1.0 + x[0] * x[1])

    @vectorised
    def rho_v(x):
        return 1.0 + x[0] * x[1]

    ref = _model(golden, is_density=rho).densities
    got = _model(golden, is_density=rho_v).densities
    assert np.array_equal(ref, got)
    with pytest.raises(TypeError, match="float array"):
        _model(golden, is_density=vectorised(lambda x: 1.0))
