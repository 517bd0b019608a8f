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


# -- boundary conditions and tips ------------------------------------------

def disp_bc(x):
    """conftest.py:24-40."""
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


@vectorised
def disp_bc_v(x):
    out = np.full((3, x.shape[1]), np.nan)
    lo = x[0] < 1.5 * 0.1
    hi = ~lo & (x[0] > 1.0 - 1.5 * 0.1)
    out[:, lo] = np.array([-1.0, 0.0, 0.0])[:, None]
    out[:, hi] = np.array([1.0, 0.0, 0.0])[:, None]
    return out


def force_bc(x):
    return [None, 1.0e-3 * (1 + x[1]), None] if x[0] > 0.95 else [None] * 3


@vectorised
def force_bc_v(x):
    out = np.full((3, x.shape[1]), np.nan)
    on = x[0] > 0.95
    out[1, on] = 1.0e-3 * (1 + x[1, on])
    return out


def tip(x):
    """Three tip groups; the x-DOF of right-hand nodes is in two."""
    t = [None, None, None]
    if x[0] > 0.9:
        t[0] = ("right", "edge")
        t[1] = "right"
    elif x[1] > 0.95:
        t[2] = "top"
        t[0] = "edge"
    return t


@vectorised
def tip_v(x):
    out = np.full((3, x.shape[1]), None, dtype=object)
    right = x[0] > 0.9
    top = ~right & (x[1] > 0.95)
    for m in np.nonzero(right)[0]:
        out[0, m] = ("right", "edge")
    out[1, right] = "right"
    out[2, top] = "top"
    out[0, top] = "edge"
    return out


def _bc_model(golden, d, f, t):
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        return Model(golden / "example_mesh.vtk", Recorder(1e-3, "cl"),
                     horizon=0.1, critical_stretch=0.005, bond_stiffness=1.0,
                     is_displacement_boundary=d, is_force_boundary=f,
                     is_tip=t)


def _bc_state(m):
    return (m.bc_types, m.bc_values, m.force_bc_types, m.force_bc_values,
            list(m.tip_types.items()), list(m.ntips.items()))


@pytest.mark.parametrize("d,f,t", [
    (disp_bc_v, force_bc_v, tip_v),
    (disp_bc_v, force_bc, tip),
    (disp_bc, force_bc_v, tip),
    (disp_bc, force_bc, tip_v),
])
def test_vectorised_boundary_conditions_equal_per_node_calls(golden, d, f, t):
    """Every mix of vectorised and per-node callbacks gives the per-node
    arrays bitwise, force densities included (value / volume / number of
    force nodes), and the same tip groups in the same order."""
    ref = _bc_state(_bc_model(golden, disp_bc, force_bc, tip))
    got = _bc_state(_bc_model(golden, d, f, t))
    for a, b in zip(ref[:4], got[:4]):
        assert a.dtype == b.dtype and np.array_equal(a, b)
    assert got[4] == ref[4] and got[5] == ref[5]
    assert np.count_nonzero(ref[2]) > 0 and len(ref[4]) == 3


def test_default_callbacks_set_nothing(golden):
    m = _bc_model(golden, None, None, None)
    assert not m.bc_types.any() and not m.force_bc_types.any()
    assert not m.bc_values.any() and not m.force_bc_values.any()
    assert m.tip_types == {} and list(m.ntips) == ["model"]


def test_vectorised_boundary_shape_errors(golden):
    @vectorised
    def wrong(x):
        return np.zeros(x.shape[1])

    @vectorised
    def wrong_tip(x):
        return [None] * x.shape[1]

    with pytest.raises(TypeError, match="is_displacement_boundary"):
        _bc_model(golden, wrong, None, None)
    with pytest.raises(TypeError, match="is_force_boundary"):
        _bc_model(golden, None, wrong, None)
    with pytest.raises(TypeError, match="is_tip"):
        _bc_model(golden, None, None, wrong_tip)
