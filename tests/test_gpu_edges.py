"""Edge cases of the step kernel against the oracle, through the C ABI: the
smallest mesh (one bond), a very ragged one (a node with a thousand bonds
beside nodes with one), a row over the LDS halo capacity (rejected cleanly),
and a step in which every bond breaks.  "compact_exact" is the Cython layout
in its exact-order mode, held to the oracle bitwise."""
import numpy as np
import pytest

import oracle
import pdcases as pc
from probabilistic_peridynamics_amd import _lib
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


def _lists(rows, layout):
    """(family, nlist, n_neigh) from per-node neighbour lists."""
    n = len(rows)
    fam = np.array([len(r) for r in rows], dtype=np.int32)
    if layout == "slot":
        k = 1 << (int(fam.max()) - 1).bit_length()
        nl = -np.ones((n, k), dtype=np.int32)
    else:
        k = int(fam.max())
        nl = np.zeros((n, k), dtype=np.int32)
    for i, r in enumerate(rows):
        nl[i, :len(r)] = r
    return fam, nl, fam.copy()


def _run(layout, coords, rows, steps, bc_types, bc_values, dmags, k=1e12,
         s0=0.01, dt=1e-9, u0=None):
    """Run the HIP integrator and the oracle side by side; returns both
    states after ``steps`` steps."""
    n = len(coords)
    fam, nl, nn = _lists(rows, layout)
    vol = np.full(n, 1e-6)
    fbt = np.zeros((n, 3), np.int32)
    fbv = np.zeros((n, 3))
    integ = (it.EulerHIP(dt) if layout == "slot" else
             it.Euler(dt, exact_order=layout == "compact_exact"))
    integ.build(n, 3, nl.shape[1], coords, vol, fam, bc_types, bc_values,
                fbt, fbv, None, None, None)
    z = np.zeros((n, 3))
    u = z if u0 is None else u0
    integ.create_buffers(nl, nn, k, s0, None, u, z, z, z, z, np.zeros(n),
                         None, 1, 1)
    if layout == "slot":
        ref = oracle.SlotStepper(0, dt, coords, vol, fam, bc_types,
                                 bc_values, fbt, fbv, nl, k, s0, u=u)
    else:
        ref = oracle.CythonEuler(dt, coords, vol, fam, bc_types, bc_values,
                                 fbt, fbv, nl, nn, k, s0, u=u)
    for s in range(steps):
        integ(dmags[s], 0.0)
        ref(dmags[s], 0.0)
    got = integ.write(None, None, None, None, None, None, None, None)
    if layout == "slot":
        rn, rd = ref.connectivity()
    else:
        rn, rd = ref.n_neigh, ref.damage()
    return got, (ref.u, ref.nlist, rn, rd)


def _check(got, ref, exact=False):
    u, _, _, _, _, dmg, nl, nn = got
    ru, rnl, rnn, rdmg = ref
    np.testing.assert_array_equal(nl, rnl)
    np.testing.assert_array_equal(nn, rnn)
    np.testing.assert_array_equal(dmg, rdmg)
    assert pc.rel_l2(u, ru) < 1e-10
    if exact:
        assert np.array_equal(u, ru)


@pytest.mark.parametrize("layout", ["slot", "compact", "compact_exact"])
def test_one_bond(layout):
    """Two nodes, one bond, one end held, the other pulled until the bond
    breaks (its stretch reaches 0.01 = s0 exactly, in rounding, at step 6)."""
    coords = np.array([[0.0, 0.0, 0.0], [1e-3, 0.0, 0.0]])
    bct = np.ones((2, 3), np.int32)
    bcv = np.zeros((2, 3))
    bcv[1, 0] = 1.0
    d = 2e-6 * np.arange(1, 21)
    got, ref = _run(layout, coords, [[1], [0]], 20, bct, bcv, d)
    _check(got, ref, layout == "compact_exact")
    assert got[7][0] == 0 and got[5][0] == 1.0


@pytest.mark.parametrize("layout", ["slot", "compact", "compact_exact"])
def test_ragged_star(layout):
    """A hub bonded to 1000 nodes on a rough shell, each of those bonded to
    the hub only: one row of 1000 next to rows of 1 (tiles cut by the halo
    capacity, idle lanes, long and short chunk rows in one tile)."""
    rng = np.random.default_rng(3)
    m = 1000
    d = rng.normal(size=(m, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    d *= rng.uniform(0.9e-3, 1.1e-3, m)[:, None]
    coords = np.vstack([[0.0, 0.0, 0.0], d])
    rows = [list(range(1, m + 1))] + [[0]] * m
    n = m + 1
    bct = np.ones((n, 3), np.int32)  # the shell is held
    bcv = np.zeros((n, 3))
    bcv[0] = [1.0, 0.5, -0.25]
    dm = 3e-6 * np.arange(1, 16)
    got, ref = _run(layout, coords, rows, 15, bct, bcv, dm)
    _check(got, ref, layout == "compact_exact")
    broken = m - got[7][0]
    assert 0 < broken < m


def test_row_over_halo_capacity_is_rejected():
    """A node with more bonds than a tile's halo holds: create_buffers fails
    with a limit error before any launch, and the library stays usable."""
    m = 1400
    rng = np.random.default_rng(1)
    coords = np.vstack([[0.0, 0.0, 0.0], rng.normal(size=(m, 3)) * 1e-3])
    rows = [list(range(1, m + 1))] + [[0]] * m
    fam, nl, nn = _lists(rows, "slot")
    n = m + 1
    integ = it.EulerHIP(1e-9)
    integ.build(n, 3, nl.shape[1], coords, np.full(n, 1e-6), fam,
                np.zeros((n, 3), np.int32), np.zeros((n, 3)),
                np.zeros((n, 3), np.int32), np.zeros((n, 3)), None, None,
                None)
    z = np.zeros((n, 3))
    with pytest.raises(_lib.PDError, match="halo capacity"):
        integ.create_buffers(nl, nn, 1e12, 0.01, None, z, z, z, z, z,
                             np.zeros(n), None, 1, 1)
    with pytest.raises(_lib.PDError):
        integ(0.0, 0.0)
    # a valid mesh still runs afterwards
    coords2 = np.array([[0.0, 0.0, 0.0], [1e-3, 0.0, 0.0]])
    bct = np.zeros((2, 3), np.int32)
    got, ref = _run("slot", coords2, [[1], [0]], 3, bct, np.zeros((2, 3)),
                    np.zeros(3))
    _check(got, ref)


@pytest.mark.parametrize("layout", ["slot", "compact", "compact_exact"])
def test_every_bond_breaks_in_the_first_step(layout):
    """An initial displacement that stretches every bond by 50 %: every
    bond breaks in step 1, every node's damage is 1, and no force acts."""
    w_coords = pc.read_vtk_points(pc.GOLDEN / "example_mesh.vtk")
    fam, nl = _lib.neighbour_list(w_coords, 0.1)
    rows = [list(nl[i, :fam[i]]) for i in range(len(fam))]
    n = len(fam)
    u0 = 0.5 * w_coords
    bct = np.zeros((n, 3), np.int32)
    got, ref = _run(layout, w_coords, rows, 2, bct, np.zeros((n, 3)),
                    np.zeros(2), u0=u0)
    _check(got, ref, layout == "compact_exact")
    assert np.all(got[7] == 0) and np.all(got[5] == 1.0)
    assert np.all(got[3] == 0.0)
