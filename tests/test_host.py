"""CPU tests: the C ABI library, mesh reading and the Model's set-up.

No GPU calls: the Model is built with a recording integrator that checks
what Model.__init__ / simulate hand over (model.py:475-479, 1569-1571).
"""
import ctypes
import re
import warnings

import numpy as np
import pytest

import pdcases as pc
from probabilistic_peridynamics_amd import _lib, mesh, configs
from probabilistic_peridynamics_amd import integrators as it
from probabilistic_peridynamics_amd.model import (
    Model, initial_crack_helper, DimensionalityError, FamilyError,
    DamageModelError, InvalidIntegrator, create_crack)

HEADER = pc.GOLDEN.parents[1] / "include" / "pd_hip.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pd_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(str(_lib.LIB_PATH))
    names = declared_functions()
    assert len(names) >= 15
    for name in names:
        assert hasattr(L, name), name
    assert set(names) == set(_lib.SIGNATURES)


def test_library_loads_and_reports_errors():
    assert isinstance(_lib.lib().pd_last_error(), bytes)
    n = _lib.device_count()
    assert n >= 0
    # a rejected call leaves its message for pd_last_error
    rc = _lib.lib().pd_set_imprecise_surface_correction(
        -1, 0, 0, None, None, None, 1.0, 1.0)
    assert rc == _lib.PD_E_ARG
    assert b"required" in _lib.lib().pd_last_error()


def test_neighbour_list_matches_reference_sets(golden):
    """Same neighbour sets as expected_connectivity_basic(_cl).npz (the
    reference's KD-tree, model.py:564-566); rows ascending."""
    coords = pc.read_vtk_points(golden / "example_mesh.vtk")
    fam, nl = _lib.neighbour_list(coords, pc.HORIZON)
    z = np.load(golden / "expected_connectivity_basic.npz")
    np.testing.assert_array_equal(fam, z["n_neigh"])
    assert nl.shape == z["nlist"].shape
    for i in range(len(fam)):
        assert np.array_equal(nl[i, :fam[i]], np.sort(z["nlist"][i, :fam[i]]))
        assert np.all(nl[i, fam[i]:] == 0)


def test_neighbour_list_brute_force():
    rng = np.random.default_rng(3)
    x = rng.random((400, 3))
    h = 0.21
    fam, nl = _lib.neighbour_list(x, h, 100, -1)
    d = x[None, :, :] - x[:, None, :]
    q = (d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]
    for i in range(len(x)):
        want = np.nonzero((q[i] <= h * h) & (np.arange(len(x)) != i))[0]
        np.testing.assert_array_equal(nl[i, :fam[i]], want)
        assert np.all(nl[i, fam[i]:] == -1)


def test_vtk_reader(golden):
    pts, cells = mesh.read(golden / "example_mesh.vtk")
    np.testing.assert_array_equal(pts, pc.read_vtk_points(
        golden / "example_mesh.vtk"))
    assert cells["triangle"].shape == (4096, 3)
    assert cells["line"].shape == (128, 2)


def test_vtk_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    pts = rng.random((10, 3))
    tri = np.array([[0, 1, 2], [2, 3, 4]])
    for binary in (True, False):
        f = tmp_path / f"m{binary}.vtk"
        mesh.write_vtk(f, pts, [("triangle", tri)],
                       {"damage": rng.random(10)}, binary=binary)
        p2, c2 = mesh.read(f)
        np.testing.assert_array_equal(p2, pts)
        np.testing.assert_array_equal(c2["triangle"], tri)


class Recorder(it.Integrator):
    """An integrator that records what the Model hands it (no GPU)."""

    def __init__(self, dt, context="cython"):
        self.dt = dt
        self.context = None if context == "cython" else it.HIPContext(0)
        self._sim = None
        self.damping = 0.0

    def __call__(self, d, f):
        self.calls.append((d, f))

    def _build_special(self):
        pass

    def _create_special_buffers(self):
        pass

    def build(self, *args):
        self.build_args = args
        self.calls = []

    def create_buffers(self, *args):
        self.buffer_args = args

    def write(self, u, ud, udd, force, body_force, damage, nlist, n_neigh):
        return (u, ud, udd, force, body_force, damage, nlist, n_neigh)


def _model(golden, layout="cython", **kw):
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        return Model(golden / "example_mesh.vtk", Recorder(1e-3, layout),
                     horizon=0.1, critical_stretch=0.005,
                     bond_stiffness=18.0 * 0.05 / (np.pi * 0.1 ** 4), **kw)


def test_model_volume(golden):
    """test_model.py:261-264."""
    m = _model(golden)
    assert np.allclose(m.volume, np.load(golden / "expected_volume.npy"))


@pytest.mark.parametrize("layout,name,width", [
    ("cython", "expected_connectivity_crack.npz", 68),
    ("cl", "expected_connectivity_crack_cl.npz", 128)])
def test_model_crack_connectivity(golden, layout, name, width):
    """The cracked connectivity of conftest.py's model (test_model.py:339-361)
    holds the same bond sets as the fixtures, in the layout of the
    integrator's context (model.py:575-582)."""
    m = _model(golden, layout,
               initial_crack=initial_crack_helper(pc.is_crack),
               is_displacement_boundary=pc.simple_displacement_boundary)
    nl, nn = m.initial_connectivity
    z = np.load(golden / name)
    assert nl.shape == (2113, width)
    np.testing.assert_array_equal(nn, z["n_neigh"])
    for i in range(2113):
        assert set(nl[i, :nn[i]]) == set(z["nlist"][i, :nn[i]])
    np.testing.assert_array_equal(m.family, np.load(
        golden / "expected_connectivity_basic.npz")["n_neigh"])


def test_model_boundary_conditions(golden):
    """test_model.py:1424-1489."""
    def disp(x):
        bnd = [None, None, None]
        if x[0] < 0.1:
            bnd[0] = -1
        elif x[0] > 0.9:
            bnd[0] = 1
        return bnd

    def tip(x):
        bnd = [None, None, None]
        if x[0] < 0.1:
            bnd[0] = 1
            bnd[1] = 2
        else:
            bnd[0] = "otherwise"
            bnd[1] = 2
        return bnd

    m = _model(golden, is_displacement_boundary=disp, is_force_boundary=disp,
               is_tip=tip)
    bt, bv, fbt, fbv, _, ntips = m._set_boundary_conditions(disp, disp, tip)
    assert ntips["model"] == 2113 and ntips["1"] == 228
    assert ntips["otherwise"] == 2113 - 228 and ntips["2"] == 2113
    assert np.allclose(bt, np.load(golden / "expected_bc_types.npy"))
    assert np.allclose(bv, np.load(golden / "expected_bc_values.npy"))
    assert np.allclose(fbt, np.load(golden / "expected_force_bc_types.npy"))
    assert np.allclose(fbv, np.load(golden / "expected_force_bc_values.npy"))


def test_damage_model_plus_cs(golden):
    """test_model.py:1100-1259."""
    m = _model(golden)
    k, s, pcs, nbt, nr = m._set_damage_model(
        [[1.0, -1.0, -0.5], [1.0, 0.0, 0.0]],
        [[1.0, 1.5, 2.5], [1.0, 1000.0, 1001.0]])
    assert nr == 3 and nbt == 2
    assert np.all(pcs == np.array([[0.0, 2.0, 1.25], [0.0, 1.0, 1.0]]))
    _, _, pcs, nbt, nr = m._set_damage_model(np.array([1.0, -1.0, -0.5]),
                                             np.array([1.0, 2.0, 3.0]))
    assert np.all(pcs == [0.0, 2.0, 1.0]) and nr == 3 and nbt == 1
    _, _, pcs, nbt, nr = m._set_damage_model(np.array([[1.0], [2.0], [1.0]]),
                                             np.array([[1.0], [1.0], [1.0]]))
    assert np.all(pcs == 0) and nr == 1 and nbt == 3
    _, _, pcs, nbt, nr = m._set_damage_model([1.0], [1.0])
    assert pcs is None and nr == 1 and nbt == 1
    with pytest.raises(ValueError):
        m._set_damage_model(np.array([[1.0, 1.0, -0.5], [1.0, 0.0, 0.0]]),
                            np.array([[-1.0, 1.0, 2.5], [1.0, 1e3, 1001.0]]))
    with pytest.raises(DamageModelError):
        m._set_damage_model([1.0, 2.0], [2.0, 1.0])


def test_trilinear_tables_match_model_recurrence(golden):
    """configs.trilinear_tables == Model._set_damage_model's plus_cs."""
    m = _model(golden)
    k, s, pcs = configs.trilinear_tables(1e17)
    _, _, want, _, _ = m._set_damage_model(k, s)
    np.testing.assert_array_equal(pcs, want)
    # force continuity at the breakpoints: s*k_r + c_r == s*k_{r+1} + c_{r+1}
    for r in range(2):
        a = s[:, r] * k[:, r] + pcs[:, r]
        b = s[:, r] * k[:, r + 1] + pcs[:, r + 1]
        np.testing.assert_allclose(a, b, rtol=1e-12)


def test_model_errors(golden):
    with pytest.raises(InvalidIntegrator):
        Model(golden / "example_mesh.vtk", None, horizon=0.1,
              critical_stretch=0.005, bond_stiffness=1.0)
    with pytest.raises(DimensionalityError):
        _model(golden, dimensions=4)
    with pytest.raises(FamilyError):
        Model(np.array([[0.0, 0, 0], [10.0, 0, 0]]), Recorder(1.0),
              horizon=0.1, critical_stretch=0.005, bond_stiffness=1.0,
              volume=np.ones(2))


def test_simulate_hands_state_to_integrator(golden):
    m = _model(golden, is_displacement_boundary=pc.simple_displacement_boundary)
    mags = np.linspace(1, 10, 10)
    m.simulate(10, displacement_bc_magnitudes=mags)
    assert [c[0] for c in m.integrator.calls] == list(mags)
    (nlist, n_neigh, k, s0, pcs, u, ud, udd, force, body_force, damage,
     regimes, nreg, nbt) = m.integrator.buffer_args
    assert nlist.shape == (2113, 68) and nreg == 1 and nbt == 1
    assert k == pytest.approx(18.0 * 0.05 / (np.pi * 0.1 ** 4))
    b = m.integrator.build_args
    assert b[0] == 2113 and b[2] == 68
    np.testing.assert_array_equal(b[6], m.bc_types)


def test_create_crack_semantics():
    nl = np.array([[1, 2, 3, 0], [0, 2, 0, 0], [0, 1, 0, 0], [0, 0, 0, 0]],
                  dtype=np.int32)
    nn = np.array([3, 2, 2, 1], dtype=np.int32)
    create_crack([(0, 1)], nl, nn)
    np.testing.assert_array_equal(nl, [[3, 2, -1, 0], [2, -1, 0, 0],
                                       [0, 1, 0, 0], [0, 0, 0, 0]])
    np.testing.assert_array_equal(nn, [2, 1, 2, 1])


def test_bar_workload_sizes():
    """The 2560-node pre-cracked bar: nodes, bonds after the crack, crack
    bonds and family range."""
    w = configs.bar_euler()
    assert w.nnodes == 2560 and w.live_bonds() == 219040
    assert w.extra["crack_bonds"] == 1428
    assert (w.family.min(), w.family.max()) == (28, 122)


def test_integrator_contexts():
    """Euler keeps context None (Cython layout); the HIP integrators
    reject a non-HIPContext (integrators.py:46-50)."""
    assert it.Euler(1.0).context is None
    with pytest.raises(TypeError):
        it.EulerHIP(1.0, context=5)
    with pytest.raises(TypeError):
        it.Integrator(1.0)


def test_build_rejects_bad_arguments_before_touching_a_device():
    """pd_build's argument checks (family > 0, width, densities for the
    integrator) come before any device call, so they hold without a GPU."""
    import ctypes
    n, k = 3, 4
    coords = np.zeros((n, 3))
    vol = np.ones(n)
    bct = np.zeros((n, 3), np.int32)
    bcv = np.zeros((n, 3))

    def build(family, kk=k, integrator=0, dens=None, layout=1):
        sim = ctypes.c_void_p()
        return _lib.lib().pd_build(
            ctypes.byref(sim), 0, integrator, layout, 1e-3, 0.0, n, kk,
            _lib.dptr(coords), _lib.dptr(vol), _lib.iptr(family),
            _lib.iptr(bct), _lib.dptr(bcv), _lib.iptr(bct), _lib.dptr(bcv),
            None, None, _lib.dptr(dens))

    fam = np.array([1, 0, 2], np.int32)
    assert build(fam) == _lib.PD_E_ARG
    assert b"family" in _lib.lib().pd_last_error()
    ok = np.array([1, 1, 2], np.int32)
    assert build(ok, kk=0) == _lib.PD_E_LIMIT
    assert build(ok, dens=np.ones((n, 3))) == _lib.PD_E_ARG
    assert b"densities" in _lib.lib().pd_last_error()
    assert build(ok, integrator=2) == _lib.PD_E_ARG
    # the compact layout's rows are compacted in LDS: at most 27285 slots
    assert build(ok, kk=27286, layout=0) == _lib.PD_E_LIMIT
    assert b"27285" in _lib.lib().pd_last_error()
    assert build(ok, kk=27285, layout=0) != _lib.PD_E_LIMIT
