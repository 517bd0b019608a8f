"""HIP path against the oracle, through the C ABI (libpd_hip.so).

Bars: broken-bond set, neighbour lists (including the compact layout's
stale tail slots), regimes and damage identical; displacements within a
relative L2 error of 1e-10 (PARITY_RTOL); forces within 1e-9.
"""
import numpy as np
import pytest

import oracle
import pdcases as pc
from probabilistic_peridynamics_amd import _lib, configs
from probabilistic_peridynamics_amd import integrators as it

pytestmark = pytest.mark.gpu

PARITY_RTOL = 1e-10   # u, ud, udd: relative L2 (the path's FP64 tolerance)
FORCE_RTOL = 1e-9     # force: relative L2


def _hip_available():
    try:
        return _lib.device_count() > 0
    except OSError:
        return False


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not _hip_available():
        pytest.fail("no HIP device: the gpu tests need an MI355X")


def _state(integ):
    n, k = integ.nnodes, integ.max_neighbours
    u, ud, udd, f, bf, dmg, nl, nn = integ.write(
        None, None, None, None, None, None, None, None)
    return dict(u=u, ud=ud, udd=udd, force=f, body_force=bf, damage=dmg,
                nlist=nl, n_neigh=nn)


# --------------------------------------------------------------------------
# stretch arithmetic, bond by bond
# --------------------------------------------------------------------------

def _bond_batch(seed=0, n=20000, s0=0.005):
    rng = np.random.default_rng(seed)
    r0i = rng.uniform(-1, 1, (n, 3))
    d0 = rng.normal(size=(n, 3))
    d0 *= (10.0 ** rng.uniform(-4, 1, n))[:, None] / np.linalg.norm(
        d0, axis=1)[:, None]
    r0j = r0i + d0
    # stretches from compression to far past s0, a third placed within a
    # few ulps of +-s0 (the break decision's hardest cases)
    target = rng.uniform(-0.02, 0.02, n)
    near = rng.random(n) < 0.33
    target[near] = np.sign(rng.normal(size=near.sum())) * s0 * (
        1 + rng.integers(-8, 9, near.sum()) * 2.0 ** -52)
    ri = r0i + rng.normal(scale=1e-3, size=(n, 3))
    rj = ri + d0 * (1 + target)[:, None]
    return ri, rj, r0i, r0j


def test_stretch_bitwise_vs_oracle():
    s0 = 0.005
    ri, rj, r0i, r0j = _bond_batch()
    # special cases: zero-length current bond, NaN, equal to r0
    ri[0] = rj[0]
    rj[1, 0] = np.nan
    ri[2], rj[2] = r0i[2], r0j[2]
    got = _lib.bond_check(ri, rj, r0i, r0j, s0)
    L = oracle.lib()
    want_l = np.array([oracle.euclid(a, b) for a, b in zip(ri, rj)])
    want_l0 = np.array([oracle.euclid(a, b) for a, b in zip(r0i, r0j)])
    want_s = np.array([oracle.strain(a, b, c, d)
                       for a, b, c, d in zip(ri, rj, r0i, r0j)])
    assert L is not None
    np.testing.assert_array_equal(got["l"], want_l)
    np.testing.assert_array_equal(got["l0"], want_l0)
    np.testing.assert_array_equal(got["s_exact"], want_s)
    want_broken = ~(np.abs(want_s) < s0)
    np.testing.assert_array_equal(got["broken"].astype(bool), want_broken)
    ok = np.isfinite(want_s)
    rel = np.abs(got["s"][ok] - want_s[ok]) / np.maximum(np.abs(want_s[ok]),
                                                         1e-300)
    assert rel.max() < 1e-12
    # tension-only criterion (cl/peridynamics.cl:78)
    got_t = _lib.bond_check(ri, rj, r0i, r0j, s0, tension_only=True)
    np.testing.assert_array_equal(got_t["broken"].astype(bool),
                                  ~(want_s < s0))


def test_sqrt_core_range_edges():
    """q at and around the core's range limits and at huge/tiny scales."""
    base = np.array([0.0, 0.0, 0.0])
    qs = []
    for e in (-600, -383, -382, -300, -20, 0, 20, 300, 500, 510):
        for m in (1.0, 1.0000000000000002, 1.9999999999999998):
            qs.append(m * 2.0 ** e)
    n = len(qs)
    rj = np.zeros((n, 3))
    rj[:, 0] = np.sqrt(np.array(qs))
    rj[:, 1] = rj[:, 0] * 1e-3
    ri = np.zeros((n, 3))
    got = _lib.bond_check(ri, rj, ri, rj, 0.5)
    want = np.array([oracle.euclid(base, r) for r in rj])
    np.testing.assert_array_equal(got["l"], want)


# --------------------------------------------------------------------------
# the reference's 10-step regression model
# --------------------------------------------------------------------------

def _run_regression(cls, layout, golden, **kw):
    a = pc.regression_inputs(layout)
    n = a["coords"].shape[0]
    if cls is it.Euler:
        integ = it.Euler(pc.DT)
        dens = None
    elif cls is it.EulerHIP:
        integ = it.EulerHIP(pc.DT)
        dens = None
    else:
        integ = cls(0.0, pc.DT)
        dens = np.ones((n, 3))
    integ.build(n, 3, a["nlist"].shape[1], a["coords"], a["volume"],
                a["family"], a["bc_types"], a["bc_values"],
                a["force_bc_types"], a["force_bc_values"], None, None, dens)
    z = np.zeros((n, 3))
    integ.create_buffers(a["nlist"], a["n_neigh"], pc.BOND_STIFFNESS,
                         pc.CRITICAL_STRETCH, None, z, z, z, z, z,
                         np.zeros(n), None, 1, 1)
    for s in range(pc.STEPS):
        integ(pc.DISP_MAGS[s], 0.0)
    return integ, a, _state(integ)


def test_regression_cython_euler(golden):
    """test_regression.py:36-79 / test_integrators.py:147-196 on the GPU,
    and the oracle at the path's tolerance."""
    integ, a, got = _run_regression(it.Euler, "cython", golden)
    assert np.allclose(got["u"], np.load(golden / "expected_displacements.npy"))
    assert np.allclose(got["force"], np.load(golden / "expected_force.npy"))
    np.testing.assert_array_equal(got["damage"],
                                  np.load(golden / "expected_damage.npy"))
    z = np.load(golden / "expected_connectivity_crack.npz")
    np.testing.assert_array_equal(got["nlist"], z["nlist"])
    np.testing.assert_array_equal(got["n_neigh"], z["n_neigh"])
    ref = oracle.CythonEuler(pc.DT, a["coords"], a["volume"], a["family"],
                             a["bc_types"], a["bc_values"],
                             a["force_bc_types"], a["force_bc_values"],
                             a["nlist"], a["n_neigh"], pc.BOND_STIFFNESS,
                             pc.CRITICAL_STRETCH)
    for s in range(pc.STEPS):
        ref(pc.DISP_MAGS[s], 0.0)
    assert pc.rel_l2(got["u"], ref.u) < PARITY_RTOL
    assert pc.rel_l2(got["force"], ref.force) < FORCE_RTOL


@pytest.mark.parametrize("cls,kind,prefix", [
    (it.EulerHIP, oracle.EULER, ""),
    (it.EulerCromerHIP, oracle.EULER_CROMER, "_euler_cromer"),
    (it.VelocityVerletHIP, oracle.VELOCITY_VERLET, "_velocity_verlet"),
])
def test_regression_slot_integrators(golden, cls, kind, prefix):
    """test_integrators.py:404-455, 546-597, 688-744 on the GPU."""
    integ, a, got = _run_regression(cls, "cl", golden)
    n = a["coords"].shape[0]
    assert np.allclose(got["u"], np.load(
        golden / f"expected_displacements{prefix}.npy"))
    assert np.allclose(got["force"], np.load(
        golden / f"expected_force{prefix}.npy"))
    assert np.allclose(got["damage"], np.load(
        golden / f"expected_damage{prefix}.npy"))
    z = np.load(golden / "expected_connectivity_crack_cl.npz")
    np.testing.assert_array_equal(got["nlist"], z["nlist"])
    np.testing.assert_array_equal(got["n_neigh"], z["n_neigh"])
    ref = oracle.SlotStepper(
        kind, pc.DT, a["coords"], a["volume"], a["family"], a["bc_types"],
        a["bc_values"], a["force_bc_types"], a["force_bc_values"],
        a["nlist"], pc.BOND_STIFFNESS, pc.CRITICAL_STRETCH, damping=0.0,
        densities=None if kind == oracle.EULER else np.ones((n, 3)))
    for s in range(pc.STEPS):
        ref(pc.DISP_MAGS[s], 0.0)
    for key in ("u", "ud", "udd"):
        assert pc.rel_l2(got[key], getattr(ref, key)) < PARITY_RTOL, key
    assert pc.rel_l2(got["force"], ref.force) < FORCE_RTOL
    np.testing.assert_array_equal(got["nlist"], ref.nlist)


# --------------------------------------------------------------------------
# dynamic breaking (no reference fixture breaks a bond dynamically)
# --------------------------------------------------------------------------

def _bar(layout, pull=1e-5, dt=5e-15):
    return configs.bar_euler(layout=layout, dt=dt, pull=pull)


def test_dynamic_breaking_compact_layout():
    """Cython semantics end to end: the compacted nlist (stale tails
    included), n_neigh and damage after every checkpoint equal the
    oracle's; u within tolerance.  ~5.5k directed bonds break."""
    w = _bar("compact")
    integ = it.Euler(w.dt)
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
    d, f = w.magnitudes(60)
    broke = 0
    for s in range(60):
        integ(d[s], f[s])
        ref(d[s], f[s])
        if s % 15 == 14:
            got = _state(integ)
            np.testing.assert_array_equal(got["nlist"], ref.nlist)
            np.testing.assert_array_equal(got["n_neigh"], ref.n_neigh)
            np.testing.assert_array_equal(got["damage"], ref.damage())
            assert pc.rel_l2(got["u"], ref.u) < PARITY_RTOL
            assert pc.rel_l2(got["force"], ref.force) < FORCE_RTOL
            broke = int(w.n_neigh.sum() - ref.n_neigh.sum())
    assert broke > 4000
    assert integ.live_bonds() == int(ref.n_neigh.sum())


def _with_repeats(w, nrep=300, seed=7):
    """The compact-layout bar with ``nrep`` bonds listed twice in both of
    their rows (the second copy appended), family counting the copies."""
    rng = np.random.default_rng(seed)
    n, k = w.nlist.shape
    nl = np.zeros((n, k + 4), dtype=w.nlist.dtype)
    nl[:, :k] = w.nlist
    nn = w.n_neigh.copy()
    done = 0
    for i in rng.permutation(n):
        if done == nrep:
            break
        if nn[i] == 0 or nn[i] >= k + 4:
            continue
        j = int(nl[i, rng.integers(nn[i])])
        if nn[j] >= k + 4 or i in nl[j, w.n_neigh[j]:nn[j]] or \
                j in nl[i, w.n_neigh[i]:nn[i]]:
            continue
        nl[i, nn[i]] = j
        nn[i] += 1
        nl[j, nn[j]] = i
        nn[j] += 1
        done += 1
    w.nlist, w.n_neigh, w.family = nl, nn, nn.copy()
    return w


def test_dynamic_breaking_compact_layout_with_repeats():
    """Rows that list a bond twice: the reference's compaction removes the
    first position holding a broken bond's key each time it breaks; lists,
    counts and damage equal the oracle's at every checkpoint."""
    w = _with_repeats(_bar("compact"))
    integ = it.Euler(w.dt, exact_order=False)
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
    d, f = w.magnitudes(60)
    for s in range(60):
        integ(d[s], f[s])
        ref(d[s], f[s])
        if s % 15 == 14:
            got = _state(integ)
            np.testing.assert_array_equal(got["nlist"], ref.nlist)
            np.testing.assert_array_equal(got["n_neigh"], ref.n_neigh)
            np.testing.assert_array_equal(got["damage"], ref.damage())
            assert pc.rel_l2(got["u"], ref.u) < PARITY_RTOL
    assert int(w.n_neigh.sum() - ref.n_neigh.sum()) > 4000


@pytest.mark.parametrize("tension_only", [False, True])
def test_dynamic_breaking_slot_layout(tension_only):
    w = _bar("slot", pull=1e-5)
    integ = configs.make_integrator(w)
    integ.set_break_mode(tension_only)
    ref = oracle.SlotStepper(0, w.dt, w.coords, w.volume, w.family,
                             w.bc_types, w.bc_values, w.force_bc_types,
                             w.force_bc_values, w.nlist, w.bond_stiffness,
                             w.critical_stretch, tension_only=tension_only)
    d, f = w.magnitudes(60)
    for s in range(60):
        integ(d[s], f[s])
        ref(d[s], f[s])
    got = _state(integ)
    nn, dmg = ref.connectivity()
    np.testing.assert_array_equal(got["nlist"], ref.nlist)
    np.testing.assert_array_equal(got["n_neigh"], nn)
    np.testing.assert_array_equal(got["damage"], dmg)
    assert pc.rel_l2(got["u"], ref.u) < PARITY_RTOL
    assert int(w.n_neigh.sum() - nn.sum()) > 1000


def _composite(seed=5):
    dx = 0.01
    h = 3.01 * dx
    coords = configs.lattice((24, 10, 8), dx, jitter=0.2, seed=seed,
                             origin=(0, -0.05, 0))
    fam, nl, nn = configs.neighbours(coords, h, "slot")
    n = len(coords)
    k, s, pcs = configs.trilinear_tables(configs.micromodulus(h))
    side = (coords[:, 1] >= 0).astype(np.int32)
    jj = np.where(nl >= 0, nl, 0)
    bt = np.where(side[:, None] == side[jj], side[:, None], 2).astype(
        np.int32)
    bt[nl < 0] = 0
    bct = np.zeros((n, 3), np.int32)
    bcv = np.zeros((n, 3))
    left = coords[:, 0] < h
    right = coords[:, 0] > 24 * dx - h
    bct[left | right, 0] = 1
    bcv[left, 0] = -1
    bcv[right, 0] = 1
    return configs.Workload(
        name="composite_small", coords=coords, volume=np.full(n, dx ** 3),
        family=fam, nlist=nl, n_neigh=nn, horizon=h, bond_stiffness=k,
        critical_stretch=s, plus_cs=pcs, nregimes=3, nbond_types=3,
        bond_types=bt, dt=1e-6, integrator="euler_cromer", damping=2.5e6,
        densities=np.full((n, 3), 2400.0), bc_types=bct, bc_values=bcv,
        force_bc_types=np.zeros((n, 3), np.int32),
        force_bc_values=np.zeros((n, 3)), disp_rate=4e-6)


def test_nlinear_composite_regimes_and_breaks():
    """bond_force3 semantics: regimes walk up/down one per step, bonds
    break past the last threshold; regimes, nlist, damage exact."""
    w = _composite()
    integ = configs.make_integrator(w)
    ref = oracle.SlotStepper(
        1, w.dt, w.coords, w.volume, w.family, w.bc_types, w.bc_values,
        w.force_bc_types, w.force_bc_values, w.nlist, w.bond_stiffness,
        w.critical_stretch, damping=w.damping, densities=w.densities,
        bond_types=w.bond_types, plus_cs=w.plus_cs, nregimes=3)
    d, f = w.magnitudes(160)
    for s in range(160):
        integ(d[s], f[s])
        ref(d[s], f[s])
        if s % 40 == 39:
            got = _state(integ)
            np.testing.assert_array_equal(got["nlist"], ref.nlist)
            live = ref.nlist >= 0
            np.testing.assert_array_equal(integ.regimes()[w.nlist >= 0],
                                          ref.regimes[w.nlist >= 0])
            for key in ("u", "ud", "udd"):
                assert pc.rel_l2(got[key], getattr(ref, key)) < PARITY_RTOL
            assert live.sum() == integ.live_bonds()
    nn, dmg = ref.connectivity()
    got = _state(integ)
    np.testing.assert_array_equal(got["damage"], dmg)
    assert (ref.regimes == 2).sum() > 1000
    assert int(w.n_neigh.sum() - nn.sum()) > 500


def test_stiffness_corrections_and_node_field():
    """bond_force2 (per-bond corrections) and a per-node stiffness field
    k_ij = (c_i + c_j)/2, with breaking."""
    w = _bar("slot", pull=1e-5)
    rng = np.random.default_rng(7)
    corr = np.where(w.nlist >= 0, rng.uniform(0.5, 2.0, w.nlist.shape), 0.0)
    cnode = w.bond_stiffness * np.exp(0.3 * rng.normal(size=w.nnodes))
    integ = it.EulerHIP(w.dt)
    integ.build(w.nnodes, 3, w.max_neighbours, w.coords, w.volume, w.family,
                w.bc_types, w.bc_values, w.force_bc_types,
                w.force_bc_values, corr, None, None)
    z = np.zeros((w.nnodes, 3))
    integ.create_buffers(w.nlist, w.n_neigh, w.bond_stiffness,
                         w.critical_stretch, None, z, z, z, z, z,
                         np.zeros(w.nnodes), None, 1, 1)
    integ.set_node_stiffness(cnode)
    ref = oracle.SlotStepper(0, w.dt, w.coords, w.volume, w.family,
                             w.bc_types, w.bc_values, w.force_bc_types,
                             w.force_bc_values, w.nlist, w.bond_stiffness,
                             w.critical_stretch, stiffness_corrections=corr,
                             cnode=cnode)
    d, f = w.magnitudes(50)
    for s in range(50):
        integ(d[s], f[s])
        ref(d[s], f[s])
    got = _state(integ)
    np.testing.assert_array_equal(got["nlist"], ref.nlist)
    assert pc.rel_l2(got["u"], ref.u) < PARITY_RTOL
    assert pc.rel_l2(got["force"], ref.force) < FORCE_RTOL


@pytest.mark.parametrize("lpn,tile", [(1, 256), (2, 128), (4, 256), (8, 64),
                                      (4, 64)])
def test_launch_shapes_agree(lpn, tile):
    """Every lanes-per-node / tile size gives the oracle's result."""
    w = _bar("slot", pull=1e-5)
    integ = configs.make_integrator(w, lanes_per_node=lpn, tile_nodes=tile)
    st = integ.tile_stats()
    assert st["ntiles"] >= w.nnodes // tile
    ref = oracle.SlotStepper(0, w.dt, w.coords, w.volume, w.family,
                             w.bc_types, w.bc_values, w.force_bc_types,
                             w.force_bc_values, w.nlist, w.bond_stiffness,
                             w.critical_stretch)
    d, f = w.magnitudes(40)
    integ.run(40, d, f)
    for s in range(40):
        ref(d[s], f[s])
    got = _state(integ)
    np.testing.assert_array_equal(got["nlist"], ref.nlist)
    assert pc.rel_l2(got["u"], ref.u) < PARITY_RTOL


def test_nan_displacement_breaks_every_bond_of_the_node():
    """NaN stretch breaks (!(|s| < s0) is true for NaN), both layouts."""
    w = _bar("compact")
    u0 = np.zeros((w.nnodes, 3))
    u0[1000, 1] = np.nan
    integ = it.Euler(w.dt)
    integ.build(w.nnodes, 3, w.max_neighbours, w.coords, w.volume, w.family,
                w.bc_types, w.bc_values, w.force_bc_types,
                w.force_bc_values, None, None, None)
    integ.create_buffers(w.nlist, w.n_neigh, w.bond_stiffness,
                         w.critical_stretch, None, u0, u0 * 0, u0 * 0,
                         u0 * 0, u0 * 0, np.zeros(w.nnodes), None, 1, 1)
    integ(0.0, 0.0)
    got = _state(integ)
    ref = oracle.CythonEuler(w.dt, w.coords, w.volume, w.family, w.bc_types,
                             w.bc_values, w.force_bc_types,
                             w.force_bc_values, w.nlist, w.n_neigh,
                             w.bond_stiffness, w.critical_stretch, u=u0)
    ref(0.0, 0.0)
    assert got["n_neigh"][1000] == 0
    np.testing.assert_array_equal(got["nlist"], ref.nlist)
    np.testing.assert_array_equal(got["n_neigh"], ref.n_neigh)


def test_restart_is_bitwise_and_stateless(golden):
    """test_model.py:1530-1563 and :1628-1670: repeated runs and a restart
    from a written state give identical results, bit for bit."""
    a = pc.regression_inputs("cl")
    n = a["coords"].shape[0]

    def fresh():
        integ = it.EulerHIP(pc.DT)
        integ.build(n, 3, 128, a["coords"], a["volume"], a["family"],
                    a["bc_types"], a["bc_values"], a["force_bc_types"],
                    a["force_bc_values"], None, None, None)
        return integ

    mags = np.linspace(0, 0.00001 * 99 / 2, 100)
    z = np.zeros((n, 3))
    one = fresh()
    one.create_buffers(a["nlist"], a["n_neigh"], pc.BOND_STIFFNESS,
                       pc.CRITICAL_STRETCH, None, z, z, z, z, z,
                       np.zeros(n), None, 1, 1)
    one(mags[0], 0.0)
    s1 = _state(one)
    one.create_buffers(s1["nlist"], s1["n_neigh"], pc.BOND_STIFFNESS,
                       pc.CRITICAL_STRETCH, None, s1["u"], s1["ud"], z, z, z,
                       np.zeros(n), None, 1, 1)
    one(mags[1], 0.0)
    r1 = _state(one)
    two = fresh()
    for _ in range(2):
        two.create_buffers(a["nlist"], a["n_neigh"], pc.BOND_STIFFNESS,
                           pc.CRITICAL_STRETCH, None, z, z, z, z, z,
                           np.zeros(n), None, 1, 1)
        two(mags[0], 0.0)
        two(mags[1], 0.0)
        r2 = _state(two)
        for key in ("u", "force", "damage", "nlist", "n_neigh"):
            np.testing.assert_array_equal(r1[key], r2[key])


# --------------------------------------------------------------------------
# full-size workloads: size-independent properties
# --------------------------------------------------------------------------

def _momentum_defect(st, volume):
    """|sum_i V_i F_i| / sum_i |V_i F_i| over the bond forces (body_force):
    each bond adds f V_j to i and -f V_i to j, so the sum vanishes."""
    vf = volume[:, None] * st["body_force"]
    return np.abs(vf.sum(axis=0)).max() / np.abs(vf).sum(axis=0).max()


@pytest.mark.parametrize("name", ["beam_vv", "composite_ec"])
def test_full_size_momentum_and_restart(name):
    w = getattr(configs, name)()
    integ = configs.make_integrator(w)
    d, f = w.magnitudes(12)
    integ.run(6, d[:6], f[:6])
    mid = _state(integ)
    defect = _momentum_defect(mid, w.volume)
    print(f"{name}: momentum defect {defect:.3e}")
    assert defect < 1e-13
    integ.run(6, d[6:], f[6:])
    end = _state(integ)
    # restarting from the mid-run state reproduces the second half bitwise
    again = configs.make_integrator(w)
    again.create_buffers(mid["nlist"], mid["n_neigh"], w.bond_stiffness,
                         w.critical_stretch, w.plus_cs, mid["u"], mid["ud"],
                         mid["udd"], None, None, None,
                         integ.regimes() if w.nregimes > 1 else None,
                         w.nregimes, w.nbond_types)
    again.run(6, d[6:], f[6:])
    st = _state(again)
    np.testing.assert_array_equal(st["u"], end["u"])
    np.testing.assert_array_equal(st["nlist"], end["nlist"])
    # another launch shape sums each node's bonds in another order only
    other = configs.make_integrator(w, lanes_per_node=2, tile_nodes=256)
    other.run(12, d, f)
    assert pc.rel_l2(_state(other)["u"], end["u"]) < 1e-12


def _bar_sensitivity(w, steps=20):
    """Relative L2 change of the oracle's own u after ``steps`` steps when
    every free displacement is moved by one ulp, up or down at random,
    after step 2: how far the run's own rounding determines u."""
    def run(perturb):
        ref = oracle.SlotStepper(0, w.dt, w.coords, w.volume, w.family,
                                 w.bc_types, w.bc_values, w.force_bc_types,
                                 w.force_bc_values, w.nlist,
                                 w.bond_stiffness, w.critical_stretch)
        d, f = w.magnitudes(steps)
        for s in range(steps):
            ref(d[s], f[s])
            if s == 1 and perturb:
                up = np.random.default_rng(0).random(ref.u.shape) < 0.5
                free = w.bc_types == 0
                ref.u[free & up] = np.nextafter(ref.u[free & up], np.inf)
                ref.u[free & ~up] = np.nextafter(ref.u[free & ~up],
                                                 -np.inf)
        return ref.u
    return pc.rel_l2(run(True), run(False))


def _slot_order_sensitivity(w, steps):
    """Relative L2 difference of the oracle's own u after ``steps`` steps
    when every row's slots are summed in reverse order.  The OpenCL kernels
    sum a node's slots by a tree (cl/peridynamics.cl:104-115), so any slot
    order is an equally valid rounding of the reference's step: this is how
    far its u is defined on the run."""
    def run(nl):
        ref = oracle.SlotStepper(0, w.dt, w.coords, w.volume, w.family,
                                 w.bc_types, w.bc_values, w.force_bc_types,
                                 w.force_bc_values, nl, w.bond_stiffness,
                                 w.critical_stretch)
        d, f = w.magnitudes(steps)
        for s in range(steps):
            ref(d[s], f[s])
        return ref.u
    return pc.rel_l2(run(np.ascontiguousarray(w.nlist[:, ::-1])),
                     run(w.nlist))


@pytest.mark.parametrize("layout,exact", [("compact", None),
                                          ("compact", False),
                                          ("slot", None)])
def test_bar_workload_full_length(layout, exact):
    """The 2560-node pre-cracked bar exactly as specified (Euler, dt = 1e-5,
    1000 steps, pull 1e-8 per step).  The step is far past the stability
    limit: the pull throws the free end metres in two steps and a cascade
    breaks 27,328 bonds in step 3, after which the state stays put.  Lists,
    counts and damage equal the oracle's exactly in every mode.

    u: the Cython-layout Euler as constructed by default, Euler(dt), sums in
    peridynamics.pyx's order and equals the oracle bitwise.  Its fast order
    (exact_order=False) sums each node's bonds in its own order: held to 50
    times the run's measured sensitivity to one ulp of u after step 2.  The
    slot layout (the OpenCL semantics, whose kernel sums a node's slots by a
    tree) is held to 5 times the oracle's own difference under a reversed
    slot order: the run amplifies last-bit differences of the first steps'
    forces by ~1e8 per step, so u is defined only to that on this run."""
    w = configs.bar_euler(layout=layout)
    if layout == "compact":
        integ = (it.Euler(w.dt) if exact is None
                 else it.Euler(w.dt, exact_order=exact))
        integ.build(w.nnodes, 3, w.max_neighbours, w.coords, w.volume,
                    w.family, w.bc_types, w.bc_values, w.force_bc_types,
                    w.force_bc_values, None, None, None)
        z = np.zeros((w.nnodes, 3))
        integ.create_buffers(w.nlist, w.n_neigh, w.bond_stiffness,
                             w.critical_stretch, None, z, z, z, z, z,
                             np.zeros(w.nnodes), None, 1, 1)
        ref = oracle.CythonEuler(w.dt, w.coords, w.volume, w.family,
                                 w.bc_types, w.bc_values, w.force_bc_types,
                                 w.force_bc_values, w.nlist, w.n_neigh,
                                 w.bond_stiffness, w.critical_stretch)
    else:
        integ = configs.make_integrator(w)
        ref = oracle.SlotStepper(0, w.dt, w.coords, w.volume, w.family,
                                 w.bc_types, w.bc_values, w.force_bc_types,
                                 w.force_bc_values, w.nlist,
                                 w.bond_stiffness, w.critical_stretch)
    d, f = w.magnitudes(1000)
    for s in range(1000):
        integ(d[s], f[s])
        ref(d[s], f[s])
        if s < 4 or s == 999:
            got = _state(integ)
            if layout == "compact":
                rn, rd = ref.n_neigh, ref.damage()
            else:
                rn, rd = ref.connectivity()
            np.testing.assert_array_equal(got["nlist"], ref.nlist)
            np.testing.assert_array_equal(got["n_neigh"], rn)
            np.testing.assert_array_equal(got["damage"], rd)
    assert int(w.n_neigh.sum() - rn.sum()) == 27328
    err = pc.rel_l2(got["u"], ref.u)
    if layout == "compact" and exact is None:
        print(f"bar, compact, Euler(dt): u rel-L2 {err:.2e}")
        assert integ.exact_order
        np.testing.assert_array_equal(got["u"], ref.u)
        np.testing.assert_array_equal(got["force"], ref.force)
    elif layout == "compact":
        sens = _bar_sensitivity(configs.bar_euler(layout="slot"))
        print(f"bar, compact fast order: u rel-L2 {err:.2e}, one-ulp "
              f"sensitivity {sens:.2e}")
        assert sens > 1e-10  # the run is this ill-conditioned
        assert err < 50 * sens
    else:
        sens = _slot_order_sensitivity(w, 1000)
        print(f"bar, slot: u rel-L2 {err:.2e}, the oracle under a reversed "
              f"slot order {sens:.2e}")
        assert sens > 1e-4
        assert err < 5 * sens
