"""The oracle against the reference's 10-step regression fixtures.

peripy/test/test_regression.py:11-148 and test_integrators.py:147-196,
404-455, 546-597, 688-744 run the 2113-node plate for 10 steps and compare
with expected_*.npy / *.npz using np.allclose (connectivity exactly).  The
oracle must pass the same comparisons.
"""
import numpy as np
import pytest

import oracle
import pdcases as pc


def _cython_run():
    a = pc.regression_inputs("cython")
    st = oracle.CythonEuler(pc.DT, a["coords"], a["volume"], a["family"],
                            a["bc_types"], a["bc_values"],
                            a["force_bc_types"], a["force_bc_values"],
                            a["nlist"], a["n_neigh"], pc.BOND_STIFFNESS,
                            pc.CRITICAL_STRETCH)
    for s in range(pc.STEPS):
        st(pc.DISP_MAGS[s], 0.0)
    return st


def test_cython_euler_10_steps(golden):
    st = _cython_run()
    assert np.allclose(st.u, np.load(golden / "expected_displacements.npy"))
    assert np.allclose(st.force, np.load(golden / "expected_force.npy"))
    assert np.allclose(st.damage(), np.load(golden / "expected_damage.npy"))
    z = np.load(golden / "expected_connectivity_crack.npz")
    assert np.array_equal(st.nlist, z["nlist"])
    assert np.array_equal(st.n_neigh, z["n_neigh"])


def test_cython_euler_fixture_precision(golden):
    """How close the oracle is to the fixture, beyond np.allclose.

    The fixture was written by the Cython path on some x86-64 host; the
    oracle reproduces its arithmetic, so the remaining difference is at
    the level of double rounding only.
    """
    st = _cython_run()
    u_ref = np.load(golden / "expected_displacements.npy")
    f_ref = np.load(golden / "expected_force.npy")
    assert pc.rel_l2(st.u, u_ref) < 2e-13
    assert pc.rel_l2(st.force, f_ref) < 1e-12
    assert np.array_equal(st.damage(),
                          np.load(golden / "expected_damage.npy"))


def _cl_order_run():
    """The same 10 steps with cl/peridynamics.cl:65-88's arithmetic
    (xi = r0_j - r0_i, xi_eta = (u_j - u_i) + xi, f = s*k*V_j, c = xi_eta/y),
    evaluated in numpy: a plausible wrong restatement of the stretch."""
    a = pc.regression_inputs("cython")
    x0, V, nl, nn = a["coords"], a["volume"], a["nlist"], a["n_neigh"]
    n = len(x0)
    I = np.repeat(np.arange(n), nn)
    J = np.concatenate([nl[i, :nn[i]] for i in range(n)])
    u = np.zeros((n, 3))
    for s in range(pc.STEPS):
        xi = x0[J] - x0[I]
        xe = (u[J] - u[I]) + xi
        l0 = np.sqrt((xi[:, 0] * xi[:, 0] + xi[:, 1] * xi[:, 1])
                     + xi[:, 2] * xi[:, 2])
        y = np.sqrt((xe[:, 0] * xe[:, 0] + xe[:, 1] * xe[:, 1])
                    + xe[:, 2] * xe[:, 2])
        f = (y - l0) / l0 * pc.BOND_STIFFNESS * V[J]
        F = np.zeros((n, 3))
        np.add.at(F, I, f[:, None] * (xe / y[:, None]))
        u = np.where(a["bc_types"] == 0, u + pc.DT * F,
                     pc.DISP_MAGS[s] * a["bc_values"])
    return u, F


def test_fixture_discriminates_arithmetic_order(golden):
    """Negative control for the thresholds above: the OpenCL-order
    arithmetic passes np.allclose like the oracle, but misses the fixture by
    more than the bounds the oracle meets."""
    u, F = _cl_order_run()
    u_ref = np.load(golden / "expected_displacements.npy")
    f_ref = np.load(golden / "expected_force.npy")
    assert np.allclose(u, u_ref) and np.allclose(F, f_ref)
    assert pc.rel_l2(u, u_ref) > 2e-13 or pc.rel_l2(F, f_ref) > 1e-12


@pytest.mark.parametrize("integrator,prefix,conn", [
    (oracle.EULER, "", "expected_connectivity_crack_cl.npz"),
    (oracle.EULER_CROMER, "_euler_cromer",
     "expected_connectivity_crack_cl.npz"),
    (oracle.VELOCITY_VERLET, "_velocity_verlet",
     "expected_connectivity_crack_cl.npz"),
])
def test_slot_integrators_10_steps(golden, integrator, prefix, conn):
    """EulerCL / EulerCromerCL (damping 0, rho 1) / VelocityVerletCL
    regressions (test_integrators.py:404-455, 546-597, 688-744)."""
    a = pc.regression_inputs("cl")
    n = a["coords"].shape[0]
    dens = None if integrator == oracle.EULER else np.ones((n, 3))
    st = oracle.SlotStepper(
        integrator, pc.DT, a["coords"], a["volume"], a["family"],
        a["bc_types"], a["bc_values"], a["force_bc_types"],
        a["force_bc_values"], a["nlist"], pc.BOND_STIFFNESS,
        pc.CRITICAL_STRETCH, damping=0.0, densities=dens)
    for s in range(pc.STEPS):
        st(pc.DISP_MAGS[s], 0.0)
    n_neigh, dmg = st.connectivity()
    assert np.allclose(st.u, np.load(golden / f"expected_displacements{prefix}.npy"))
    assert np.allclose(st.force, np.load(golden / f"expected_force{prefix}.npy"))
    assert np.allclose(dmg, np.load(golden / f"expected_damage{prefix}.npy"))
    z = np.load(golden / conn)
    assert np.array_equal(st.nlist, z["nlist"])
    assert np.array_equal(n_neigh, z["n_neigh"])
    if integrator == oracle.VELOCITY_VERLET:
        assert np.allclose(st.ud, np.load(
            golden / "expected_velocities_velocity_verlet.npy"))
        assert np.allclose(st.udd, np.load(
            golden / "expected_accelerations_velocity_verlet.npy"))
    if integrator == oracle.EULER_CROMER:
        assert np.allclose(st.ud, np.load(
            golden / "expected_velocities_euler_cromer.npy"))


def test_crack_fixture_from_basic(golden):
    """initial_crack_helper + create_crack on the pre-crack list give the
    post-crack fixtures (model.py:588-595, conftest.py:43-63), both
    layouts."""
    coords = pc.read_vtk_points(golden / "example_mesh.vtk")
    basic = np.load(golden / "expected_connectivity_basic.npz")
    for layout, name in (("cython", "expected_connectivity_crack.npz"),
                         ("cl", "expected_connectivity_crack_cl.npz")):
        fam = basic["n_neigh"].astype(np.int32)
        if layout == "cython":
            nl = basic["nlist"].astype(np.int32).copy()
        else:
            nl = -np.ones((len(fam), 128), dtype=np.int32)
            for i in range(len(fam)):
                nl[i, :fam[i]] = basic["nlist"][i, :fam[i]]
        n_neigh = fam.copy()
        crack = [(i, int(j)) for i in range(len(fam))
                 for j in nl[i][:n_neigh[i]]
                 if i < j and pc.is_crack(coords[i], coords[j])]
        oracle.create_crack(crack, nl, n_neigh)
        z = np.load(golden / name)
        assert np.array_equal(nl, z["nlist"]), layout
        assert np.array_equal(n_neigh, z["n_neigh"]), layout
