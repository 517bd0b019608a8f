"""The oracle against the reference's own known-answer tests.

Each test restates the inputs and expected values of a reference test
(file:line given) and runs them through the oracle (oracle/pd_oracle.c).
"""
import numpy as np
import pytest
from scipy.spatial.distance import euclidean

import oracle


# --- spatial.pyx: peripy/test/test_spatial.py ---------------------------

def test_euclid_vs_scipy():
    """test_spatial.py:10-23 (random, and negative coordinates)."""
    rng = np.random.default_rng(0)
    for _ in range(100):
        r1 = rng.random(3)
        r2 = rng.random(3) * rng.choice([-1.0, 1.0])
        assert np.isclose(euclidean(r1, r2), oracle.euclid(r1, r2))


def test_euclid_operation_order():
    """sqrt((d0^2 + d1^2) + d2^2), not any other association.

    Chosen so that the three possible summation orders give three different
    doubles: only the reference's left-to-right order passes.
    """
    r1 = np.zeros(3)
    r2 = np.array([1.0, 1e-8, 1e-8 * (1 + 2 ** -20)])
    d = r2 - r1
    sq = d * d
    want = np.sqrt((sq[0] + sq[1]) + sq[2])
    assert oracle.euclid(r1, r2) == want


def test_strain_closed_forms():
    """test_spatial.py:42-78."""
    z = np.zeros(3)
    x1, x2, x3 = (np.array([v, 0.0, 0.0]) for v in (1.0, 2.0, 3.0))
    assert oracle.strain(z, x2, z, x1) == 1.0
    assert oracle.strain2(2.0, z, x1) == 1.0
    assert oracle.strain(z, x1, z, x1) == 0.0
    assert oracle.strain2(1.0, z, x1) == 0.0
    assert oracle.strain(z, x3, z, x2) == 0.5
    assert oracle.strain2(3.0, z, x2) == 0.5


def test_strain_vs_scipy():
    """test_spatial.py:28-41."""
    rng = np.random.default_rng(1)
    for _ in range(100):
        r10, r20 = rng.random(3), rng.random(3)
        r1 = r10 + rng.random(3) * 0.1
        r2 = r20 + rng.random(3) * 0.1
        l0 = euclidean(r10, r20)
        want = (euclidean(r1, r2) - l0) / l0
        assert np.isclose(oracle.strain(r1, r2, r10, r20), want)
        assert np.isclose(oracle.strain2(euclidean(r1, r2), r10, r20), want)


# --- peridynamics.pyx: peripy/test/test_peridynamics.py ------------------

def test_damage():
    """test_peridynamics.py:7-15."""
    family = np.array([10, 5, 5, 1, 5, 7, 10, 3, 3, 4], dtype=np.int32)
    n_neigh = np.array([5, 5, 3, 0, 4, 5, 8, 3, 2, 1], dtype=np.int32)
    got = oracle.damage(n_neigh, family)
    assert np.allclose(got, (family - n_neigh) / family)
    # (family - n)/family: exactly numpy's float64 evaluation of it
    assert np.array_equal(got, (family.astype(float) - n_neigh) / family)


R0_5 = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0],
                 [2.0, 0.0, 0.0], [0.0, 0.0, 1.0]])


def test_break_bonds():
    """test_peridynamics.py:18-57: exact nlist including stale tail slots."""
    nl = np.array([[1, 2, 4], [0, 3, 0], [0, 0, 0], [1, 0, 0], [0, 0, 0]],
                  dtype=np.int32)
    n_neigh = np.array([3, 2, 1, 1, 1], dtype=np.int32)
    r = np.array([[0.0, 0.0, 0.0], [2.0, 0.0, 0.0], [0.0, 1.0, 0.0],
                  [3.0, 0.0, 0.0], [0.0, 0.0, 2.0]])
    oracle.break_bonds(r, R0_5, nl, n_neigh, 1.0)
    assert np.array_equal(nl, [[2, 2, 4], [3, 3, 0], [0, 0, 0], [1, 0, 0],
                               [0, 0, 0]])
    assert np.array_equal(n_neigh, [1, 1, 1, 1, 0])


def test_initial_force_zero():
    """test_peridynamics.py:63-91."""
    nl = np.array([[1, 2, 4], [0, 3, 0], [0, 0, 0], [1, 0, 0], [0, 0, 0]],
                  dtype=np.int32)
    n_neigh = np.array([3, 2, 1, 1, 1], dtype=np.int32)
    f = oracle.bond_force(R0_5, R0_5, nl, n_neigh, np.ones(5), 1.0,
                          np.zeros((5, 3)), np.zeros((5, 3), np.int32), 1.0)
    assert np.array_equal(f, np.zeros((5, 3)))


FORCE_VALUE = 0.00229417  # test_peridynamics.py:126


def _three_nodes():
    r0 = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [1.0, 1.0, 0.0]])
    horizon = 1.01
    k = 18.0 * 0.05 / (np.pi * horizon ** 4)
    volume = np.full(3, 0.16666667)
    u = np.array([[0.0, 0.0, 0.0], [0.05, 0.0, 0.0], [0.05, 0.05, 0.0]])
    return r0, k, volume, u


def test_force_three_nodes():
    """test_peridynamics.py:93-133."""
    r0, k, volume, u = _three_nodes()
    nl = np.array([[1, 0], [0, 2], [2, 0]], dtype=np.int32)
    n_neigh = np.array([1, 2, 1], dtype=np.int32)
    f = oracle.bond_force(r0 + u, r0, nl, n_neigh, volume, k,
                          np.zeros((3, 3)), np.zeros((3, 3), np.int32), 1.0)
    want = np.array([[FORCE_VALUE, 0, 0], [-FORCE_VALUE, FORCE_VALUE, 0],
                     [0, -FORCE_VALUE, 0]])
    assert np.allclose(f, want)
    # closed form of one bond: (s*k/l) * (r_j - r_i) * V with s = 0.05
    s = (1.05 - 1.0) / 1.0
    assert np.isclose(f[0, 0], s * k * 0.16666667)


@pytest.mark.parametrize("bc_types,bc_values,scale,dt,want", [
    # test_peridynamics.py:139-209
    (np.zeros((3, 3)), np.zeros((3, 3)), 0.0, 1.0, [1.0, 2.0, 3.0]),
    (np.zeros((3, 3)), np.zeros((3, 3)), 0.0, 2.0, [2.0, 4.0, 6.0]),
    ([[1, 1, 0]] * 3, np.zeros((3, 3)), 1.0, 2.0, [0.0, 0.0, 6.0]),
    ([[1, 1, 0]] * 3, [[2.0, 2.0, 0.0]] * 3, 0.5, 2.0, [1.0, 1.0, 6.0]),
])
def test_update_displacement(bc_types, bc_values, scale, dt, want):
    u = np.zeros((3, 3))
    f = np.array([[1.0, 2.0, 3.0]] * 3)
    oracle.update_displacement(u, np.array(bc_values, dtype=np.float64),
                               np.array(bc_types, dtype=np.int32), f, scale,
                               dt)
    assert np.array_equal(u, np.array([want] * 3))


# --- cl/euler_cromer.cl, cl/velocity_verlet.cl --------------------------

def _update(which, u, ud, udd, bc_types, bc_values, scale, damping, dt):
    L = oracle.lib()
    force = np.array([1.0, 2.0, 3.0])
    dens = np.ones(3)
    u, ud, udd = (np.array(x, dtype=np.float64) for x in (u, ud, udd))
    bt = np.array(bc_types, dtype=np.int32)
    bv = np.array(bc_values, dtype=np.float64)
    fn = (L.oracle_update_euler_cromer if which == "ec"
          else L.oracle_update_velocity_verlet)
    fn(1, oracle._d(force), oracle._d(u), oracle._d(ud), oracle._d(udd),
       oracle._i(bt), oracle._d(bv), oracle._d(dens), scale, damping, dt)
    return u, ud, udd


F3 = [1.0, 2.0, 3.0]


@pytest.mark.parametrize("args,want", [
    # test_cl_euler_cromer.py:38-92
    ((0, 0, 0, [0, 0, 0], [0, 0, 0], 0.0, 1.0, 1.0), (F3, F3, F3)),
    # :95-149
    ((0, 0, 0, [0, 0, 0], [0, 0, 0], 0.0, 1.0, 2.0),
     ([4.0, 8.0, 12.0], [2.0, 4.0, 6.0], F3)),
    # :152-209
    ((0, 1, 0, [1, 1, 0], [0, 0, 0], 1.0, 2.0, 2.0),
     ([0.0, 0.0, 6.0], [-1.0, 1.0, 3.0], [-1.0, 0.0, 1.0])),
    # :212-269
    ((0, 0, 0, [1, 1, 0], [2.0, 2.0, 0.0], 0.5, 1.0, 2.0),
     ([1.0, 1.0, 12.0], [2.0, 4.0, 6.0], F3)),
])
def test_euler_cromer_update(args, want):
    u0, ud0, udd0, bt, bv, scale, damping, dt = args
    u, ud, udd = _update("ec", np.full(3, u0), np.full(3, ud0),
                         np.full(3, udd0), bt, bv, scale, damping, dt)
    assert np.array_equal(u, want[0])
    assert np.array_equal(ud, want[1])
    assert np.array_equal(udd, want[2])


@pytest.mark.parametrize("args,want", [
    # test_cl_velocity_verlet.py:38-92
    ((0, 0, 0, [0, 0, 0], [0, 0, 0], 0.0, 1.0, 1.0),
     (F3, [0.5, 1.0, 1.5], F3)),
    # :95-152
    ((0, 0, 1, [0, 0, 0], [0, 0, 0], 0.0, 0.0, 2.0),
     ([6.0, 10.0, 14.0], [2.0, 3.0, 4.0], F3)),
    # :155-212
    ((0, 1, 0, [1, 1, 0], [0, 0, 0], 1.0, 2.0, 2.0),
     ([0.0, 0.0, 6.0], [0.0, 1.0, 2.0], [-1.0, 0.0, 1.0])),
    # :215-272
    ((0, 0, 0, [1, 1, 0], [2.0, 2.0, 0.0], 0.5, 1.0, 2.0),
     ([1.0, 1.0, 12.0], F3, F3)),
])
def test_velocity_verlet_update(args, want):
    u0, ud0, udd0, bt, bv, scale, damping, dt = args
    u, ud, udd = _update("vv", np.full(3, u0), np.full(3, ud0),
                         np.full(3, udd0), bt, bv, scale, damping, dt)
    assert np.array_equal(u, want[0])
    assert np.array_equal(ud, want[1])
    assert np.array_equal(udd, want[2])


# --- per-slot kernels: peripy/test/test_cl_peridynamics.py ---------------

def _slot_force(r0, u, nlist, volume, k, s0, corr=None, tension_only=False):
    n = r0.shape[0]
    st = oracle.SlotStepper(
        0, 1.0, r0, volume, np.ones(n, np.int32), np.zeros((n, 3), np.int32),
        np.zeros((n, 3)), np.zeros((n, 3), np.int32), np.zeros((n, 3)),
        nlist, k, s0, stiffness_corrections=corr, tension_only=tension_only,
        u=u)
    st(0.0, 1.0)
    return st


def test_slot_force_three_nodes():
    """test_cl_peridynamics.py:40-234 (bond_force1 on the 3-node model)."""
    r0, k, volume, u = _three_nodes()
    nl = np.array([[1, -1, -1, -1], [0, 2, -1, -1], [1, -1, -1, -1]],
                  dtype=np.int32)
    st = _slot_force(r0, u, nl, volume, k, 1000.0)
    want = np.array([[FORCE_VALUE, 0, 0], [-FORCE_VALUE, FORCE_VALUE, 0],
                     [0, -FORCE_VALUE, 0]])
    assert np.allclose(st.force, want)
    assert np.array_equal(st.force, st.body_force)


def test_slot_break_and_damage():
    """test_cl_peridynamics.py:236-376: -1 marks and damage."""
    volume = np.full(5, 0.16666667)
    nl = np.array([[1, 2, 4, -1], [0, 3, -1, -1], [0, -1, -1, -1],
                   [1, -1, -1, -1], [0, -1, -1, -1]], dtype=np.int32)
    u = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, 0.0],
                  [1.0, 0.0, 0.0], [0.0, 0.0, 1.0]])
    k = 18.0 * 0.05 / (np.pi * 1.1 ** 4)
    st = _slot_force(R0_5, u, nl, volume, k, 1.0)
    st.family = np.array([3, 2, 1, 1, 1], dtype=np.int32)
    assert np.array_equal(st.nlist, [[-1, 2, -1, -1], [-1, 3, -1, -1],
                                     [0, -1, -1, -1], [1, -1, -1, -1],
                                     [-1, -1, -1, -1]])
    n_neigh, dmg = st.connectivity()
    assert np.array_equal(n_neigh, [1, 1, 1, 1, 0])
    assert np.allclose(dmg, [2. / 3, 1. / 2, 0.0, 0.0, 1.0])


def test_slot_stiffness_corrections():
    """test_cl_peridynamics.py:483-585 (bond_force2: x4 and x2 forces)."""
    r0, k, volume, u = _three_nodes()
    nl = np.array([[1, -1, -1, -1], [0, 2, -1, -1], [1, -1, -1, -1]],
                  dtype=np.int32)
    corr = np.array([[4.0, 1.0, 1.0, 1.0], [4.0, 2.0, 1.0, 1.0],
                     [2.0, 1.0, 1.0, 1.0]])
    st = _slot_force(r0, u, nl, volume, k, 1000.0, corr=corr)
    want = np.array([[4 * FORCE_VALUE, 0, 0],
                     [-4 * FORCE_VALUE, 2 * FORCE_VALUE, 0],
                     [0, -2 * FORCE_VALUE, 0]])
    assert np.allclose(st.force, want)


def test_create_crack():
    """create_crack.pyx:19-43: swap with last, -1 into the vacated slot."""
    nl = np.array([[1, 2, 3, 0], [0, 2, 0, 0], [0, 1, 0, 0], [0, 0, 0, 0]],
                  dtype=np.int32)
    n_neigh = np.array([3, 2, 2, 1], dtype=np.int32)
    oracle.create_crack([(0, 1)], nl, n_neigh)
    assert np.array_equal(nl, [[3, 2, -1, 0], [2, -1, 0, 0], [0, 1, 0, 0],
                               [0, 0, 0, 0]])
    assert np.array_equal(n_neigh, [2, 1, 2, 1])


# --- n-linear (bond_force3) state machine, pinned by closed forms ----------
# No reference test calls bond_force3/4; these pin the restatement of
# cl/peridynamics.cl:313-378 against the trilinear law it must reproduce.

TRI_CS = np.array([0.002, 0.004, 0.008])
TRI_K = np.array([1.0, 0.6, 0.1]) * 1.0e11


def _plus_cs(k, cs):
    """model.py:1042-1070 recurrence: c_0 = 0,
    c_r = c_{r-1} + k_{r-1} cs_{r-1} - k_r cs_{r-1}."""
    c = np.zeros_like(k)
    for r in range(1, len(k)):
        c[r] = c[r - 1] + k[r - 1] * cs[r - 1] - k[r] * cs[r - 1]
    return c


def _trilinear(s):
    """The piecewise-linear bond force per unit volume, by geometry."""
    k0, k1, k2 = TRI_K
    s0, s1, _ = TRI_CS
    if s < s0:
        return k0 * s
    if s < s1:
        return k0 * s0 + k1 * (s - s0)
    return k0 * s0 + k1 * (s1 - s0) + k2 * (s - s1)


def _bond_stepper(regime=0, nregimes=3, k=TRI_K, cs=TRI_CS):
    r0 = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0]])
    nl = np.array([[1, -1], [0, -1]], dtype=np.int32)
    z3 = np.zeros((2, 3))
    zi = np.zeros((2, 3), np.int32)
    return oracle.SlotStepper(
        oracle.EULER, 0.0, r0, np.array([0.5, 0.25]),
        np.array([1, 1], np.int32), zi, z3, zi, z3, nl,
        np.array(k).reshape(1, -1), np.array(cs).reshape(1, -1),
        plus_cs=_plus_cs(np.array(k), np.array(cs)).reshape(1, -1),
        nregimes=nregimes, bond_types=np.zeros((2, 2), np.int32),
        regimes=np.array([[regime, 0], [regime, 0]], np.int32))


def _stretch_to(st, s):
    """Move node 1 so that the bond's stretch is s; return the stretch the
    Cython arithmetic computes from the rounded position."""
    st.u[:] = 0.0
    st.u[1, 0] = s
    return ((1.0 + s) - 1.0) / 1.0


def test_nlinear_force_follows_trilinear_law():
    st = _bond_stepper()
    for s in np.linspace(1e-4, 0.0079, 80):
        sa = _stretch_to(st, s)
        st(0.0, 0.0)
        want = _trilinear(sa)
        # node 0 is pulled towards node 1 with force density f(s) * V_1
        assert st.force[0, 0] == pytest.approx(want * 0.25, rel=1e-12)
        assert st.force[1, 0] == pytest.approx(-want * 0.5, rel=1e-12)
        assert st.nlist[0, 0] == 1


def test_nlinear_force_is_continuous_at_breakpoints():
    for b in TRI_CS[:2]:
        lo, hi = _bond_stepper(), _bond_stepper()
        for st in (lo, hi):  # walk the regime up to just below b
            for s in np.linspace(1e-4, b * (1 - 1e-6), 40):
                _stretch_to(st, s)
                st(0.0, 0.0)
        _stretch_to(lo, b * (1 - 1e-9))
        lo(0.0, 0.0)
        _stretch_to(hi, b * (1 + 1e-9))
        hi(0.0, 0.0)
        assert hi.regimes[0, 0] == lo.regimes[0, 0] + 1
        assert hi.force[0, 0] == pytest.approx(lo.force[0, 0], rel=1e-7)


def test_nlinear_one_regime_change_per_step_and_break():
    st = _bond_stepper()
    _stretch_to(st, 0.006)            # regime 2 by stretch, but one step
    st(0.0, 0.0)
    assert st.regimes[0, 0] == 1
    k1, c1 = TRI_K[1], _plus_cs(TRI_K, TRI_CS)[1]
    sa = ((1.0 + 0.006) - 1.0)
    assert st.force[0, 0] == pytest.approx((sa * k1 + c1) * 0.25, rel=1e-12)
    st(0.0, 0.0)
    assert st.regimes[0, 0] == 2
    _stretch_to(st, 0.0085)           # past the last breakpoint: breaks
    st(0.0, 0.0)
    assert st.nlist[0, 0] == -1 and st.nlist[1, 0] == -1
    assert np.all(st.force == 0.0)
    # and unloading walks down one regime per step
    down = _bond_stepper(regime=2)
    _stretch_to(down, 0.001)
    down(0.0, 0.0)
    assert down.regimes[0, 0] == 1
    down(0.0, 0.0)
    assert down.regimes[0, 0] == 0


def test_nlinear_single_regime_equals_pmb():
    """nregimes = 1 tables give bond_force1: force s*k*V and a break at
    s >= s0 (cl/peridynamics.cl:78)."""
    k, s0 = 2.0e11, 0.005
    for s, broken in ((0.003, False), (0.0049, False), (0.0051, True)):
        tab = _bond_stepper(nregimes=1, k=[k], cs=[s0])
        pmb = _bond_stepper(nregimes=1, k=[k], cs=[s0])
        pmb = oracle.SlotStepper(
            oracle.EULER, 0.0, pmb.coords, pmb.volume, pmb.family,
            pmb.bc_types, pmb.bc_values, pmb.force_bc_types,
            pmb.force_bc_values, pmb.nlist, k, s0, tension_only=True)
        for st in (tab, pmb):
            _stretch_to(st, s)
            st(0.0, 0.0)
        assert np.array_equal(tab.force, pmb.force)
        assert (tab.nlist[0, 0] == -1) == broken == (pmb.nlist[0, 0] == -1)
