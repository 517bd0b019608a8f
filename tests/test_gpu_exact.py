"""Exact-order step of the Cython path (pd_set_exact_order): the bond forces
of each node summed in peridynamics.pyx:34-99's order, so that force and u
equal the oracle's Cython restatement bitwise -- also on the 2560-node bar
run whose u the default step can hold only to the run's own rounding
sensitivity -- with the compacted lists, counts and damage exact as in the
default step."""
import numpy as np
import pytest

import oracle
import pdcases as pc
from probabilistic_peridynamics_amd import _lib, configs
from probabilistic_peridynamics_amd import integrators as it

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    try:
        ok = _lib.device_count() > 0
    except OSError:
        ok = False
    if not ok:
        pytest.fail("no HIP device: the gpu tests need an MI355X")


def _pair(w, exact=True, **kw):
    integ = it.Euler(w.dt, exact_order=exact, **kw)
    integ.build(w.nnodes, 3, w.max_neighbours, w.coords, w.volume, w.family,
                w.bc_types, w.bc_values, w.force_bc_types,
                w.force_bc_values, None, None, None)
    z = np.zeros((w.nnodes, 3))
    integ.create_buffers(w.nlist, w.n_neigh, w.bond_stiffness,
                         w.critical_stretch, None, z, z, z, z, z,
                         np.zeros(w.nnodes), None, 1, 1)
    ref = oracle.CythonEuler(w.dt, w.coords, w.volume, w.family, w.bc_types,
                             w.bc_values, w.force_bc_types,
                             w.force_bc_values, w.nlist, w.n_neigh,
                             w.bond_stiffness, w.critical_stretch)
    return integ, ref


def _assert_bitwise(integ, ref, what=""):
    u, _, _, force, _, dmg, nl, nn = integ.write(
        None, None, None, None, None, None, None, None)
    np.testing.assert_array_equal(nl, ref.nlist, err_msg=what)
    np.testing.assert_array_equal(nn, ref.n_neigh, err_msg=what)
    np.testing.assert_array_equal(dmg, ref.damage(), err_msg=what)
    assert np.array_equal(force, ref.force), (what, pc.rel_l2(force,
                                                               ref.force))
    assert np.array_equal(u, ref.u), (what, pc.rel_l2(u, ref.u))


def test_regression_case_bitwise(golden):
    """The reference's regression case (test_regression.py:36-79), every
    step."""
    a = pc.regression_inputs("cython")
    n = a["coords"].shape[0]
    w = configs.Workload(
        name="regression", coords=a["coords"], volume=a["volume"],
        family=a["family"], nlist=a["nlist"], n_neigh=a["n_neigh"],
        horizon=pc.HORIZON, bond_stiffness=pc.BOND_STIFFNESS,
        critical_stretch=pc.CRITICAL_STRETCH, dt=pc.DT, integrator="euler",
        bc_types=a["bc_types"], bc_values=a["bc_values"],
        force_bc_types=a["force_bc_types"],
        force_bc_values=a["force_bc_values"], disp_rate=0.0)
    integ, ref = _pair(w)
    for s in range(pc.STEPS):
        integ(pc.DISP_MAGS[s], 0.0)
        ref(pc.DISP_MAGS[s], 0.0)
        _assert_bitwise(integ, ref, f"step {s}")
    assert n == 2113


def test_dynamic_breaking_bitwise():
    """The bar at a stable dt with a fast pull: ~5.5k directed bonds break
    over 60 steps; the state after every 15th step."""
    w = configs.bar_euler(layout="compact", dt=5e-15, pull=1e-5)
    integ, ref = _pair(w)
    d, f = w.magnitudes(60)
    for s in range(60):
        integ(d[s], f[s])
        ref(d[s], f[s])
        if s % 15 == 14:
            _assert_bitwise(integ, ref, f"step {s}")
    assert int(w.n_neigh.sum() - ref.n_neigh.sum()) > 4000


def test_force_boundary_and_launch_shape_bitwise():
    """Force boundary conditions on the pulled end (added after the bond
    sum, peridynamics.pyx:93-97) and smaller tiles: still bitwise."""
    w = configs.bar_euler(layout="compact", dt=5e-15, pull=1e-5)
    rng = np.random.default_rng(2)
    end = w.coords[:, 0] > w.coords[:, 0].max() - 0.01
    w.force_bc_types = np.zeros((w.nnodes, 3), dtype=np.int32)
    w.force_bc_types[end, 1] = 1
    w.force_bc_values = np.zeros((w.nnodes, 3))
    w.force_bc_values[end, 1] = rng.uniform(-1, 1, end.sum()) * 1e9
    w.bc_types[end] = 0
    w.force_scale = 1.0
    integ, ref = _pair(w)
    integ.set_launch(4, 64)
    z = np.zeros((w.nnodes, 3))
    integ.create_buffers(w.nlist, w.n_neigh, w.bond_stiffness,
                         w.critical_stretch, None, z, z, z, z, z,
                         np.zeros(w.nnodes), None, 1, 1)
    d, f = w.magnitudes(30)
    for s in range(30):
        integ(d[s], f[s])
        ref(d[s], f[s])
    _assert_bitwise(integ, ref)
    assert np.any(ref.force[end, 1] != 0.0)


def test_full_length_bar_bitwise_through_run_and_graph():
    """The 2560-node pre-cracked bar exactly as specified (dt = 1e-5, 1000
    steps): one ulp of u after step 2 changes this run's final u by ~7e-10,
    so the default step meets the oracle only to that; the exact-order step
    meets it bitwise, stepping one call at a time, in the library's loop
    and in a captured graph."""
    w = configs.bar_euler(layout="compact")
    integ, ref = _pair(w)
    d, f = w.magnitudes(1000)
    for s in range(4):
        integ(d[s], f[s])
        ref(d[s], f[s])
        _assert_bitwise(integ, ref, f"step {s}")
    integ.run(500, d[4:504], f[4:504])
    integ.run_graph(495, d[504:999], f[504:999])
    for s in range(4, 999):
        ref(d[s], f[s])
    integ(d[999], f[999])
    ref(d[999], f[999])
    _assert_bitwise(integ, ref, "step 999")
    assert int(w.n_neigh.sum() - ref.n_neigh.sum()) == 27328


def test_switching_and_rejection():
    """Off again, the default step continues from the exact state (to
    rounding); the slot layout rejects the mode."""
    w = configs.bar_euler(layout="compact", dt=5e-15, pull=1e-5)
    integ, ref = _pair(w)
    d, f = w.magnitudes(20)
    for s in range(10):
        integ(d[s], f[s])
        ref(d[s], f[s])
    _assert_bitwise(integ, ref)
    integ.set_exact_order(False)
    for s in range(10, 20):
        integ(d[s], f[s])
        ref(d[s], f[s])
    u = integ.write(None, None, None, None, None, None, None, None)[0]
    assert pc.rel_l2(u, ref.u) < 1e-10
    ws = configs.bar_euler(layout="slot")
    slot = configs.make_integrator(ws)
    with pytest.raises(ValueError, match="PD_LAYOUT_COMPACT"):
        _lib.check(_lib.lib().pd_set_exact_order(slot._sim, 1))


def test_wide_rows_limit_the_exact_order():
    """The exact step compacts rows in its fold windows (12 KiB per tile):
    rows wider than 4074 slots are rejected with PD_E_LIMIT, while the fast
    order (compacted in the halo's 80 KiB) steps them and agrees with the
    oracle."""
    w = configs.bar_euler(layout="compact", dt=5e-15, pull=1e-5)
    wide = np.zeros((w.nnodes, 4075), dtype=w.nlist.dtype)
    wide[:, :w.max_neighbours] = w.nlist
    w.nlist = wide
    with pytest.raises(_lib.PDError, match="4074"):
        _pair(w)
    integ, ref = _pair(w, exact=False)
    d, f = w.magnitudes(20)
    for s in range(20):
        integ(d[s], f[s])
        ref(d[s], f[s])
    u, _, _, _, _, dmg, nl, nn = integ.write(
        None, None, None, None, None, None, None, None)
    np.testing.assert_array_equal(nl, ref.nlist)
    np.testing.assert_array_equal(nn, ref.n_neigh)
    assert pc.rel_l2(u, ref.u) < 1e-10


def test_model_simulate_exact_order_bitwise(golden, tmp_path):
    """The whole model path of the reference's regression test
    (test_regression.py:36-79): Model with the Cython-layout Euler in
    exact-order mode, the pre-cracked example mesh, a vectorised
    displacement boundary, writes every 5 steps; u, force, the lists and
    damage equal the oracle stepped on the model's own arrays bitwise."""
    import warnings
    from probabilistic_peridynamics_amd.model import (Model,
                                                      initial_crack_helper,
                                                      vectorised)

    @vectorised
    def disp_bc(x):
        out = np.full((3, x.shape[1]), np.nan)
        lo = x[0] < 1.5 * 0.1
        hi = ~lo & (x[0] > 1.0 - 1.5 * 0.1)
        out[:, lo] = np.array([-1.0, 0.0, 0.0])[:, None]
        out[:, hi] = np.array([1.0, 0.0, 0.0])[:, None]
        return out

    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        m = Model(golden / "example_mesh.vtk",
                  it.Euler(pc.DT, exact_order=True), horizon=pc.HORIZON,
                  critical_stretch=pc.CRITICAL_STRETCH,
                  bond_stiffness=pc.BOND_STIFFNESS,
                  is_displacement_boundary=disp_bc,
                  initial_crack=initial_crack_helper(pc.is_crack))
    nl0, nn0 = (a.copy() for a in m.initial_connectivity)
    ref = oracle.CythonEuler(pc.DT, m.coords, m.volume, m.family,
                             m.bc_types, m.bc_values, m.force_bc_types,
                             m.force_bc_values, nl0, nn0, pc.BOND_STIFFNESS,
                             pc.CRITICAL_STRETCH)
    (tmp_path / "w").mkdir()
    u, damage, (nl, nn), force, _, _ = m.simulate(
        pc.STEPS, displacement_bc_magnitudes=pc.DISP_MAGS, write=5,
        write_path=tmp_path / "w")
    for s in range(pc.STEPS):
        ref(pc.DISP_MAGS[s], 0.0)
    np.testing.assert_array_equal(nl, ref.nlist)
    np.testing.assert_array_equal(nn, ref.n_neigh)
    np.testing.assert_array_equal(damage, ref.damage())
    assert np.array_equal(u, ref.u), pc.rel_l2(u, ref.u)
    assert np.array_equal(force, ref.force), pc.rel_l2(force, ref.force)
