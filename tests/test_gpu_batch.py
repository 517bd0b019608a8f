"""Batched UQ samples on the GPU: every sample of a batch against the
oracle, against its own solo run (bitwise), graph runs against plain
launches, and the autotuner's state handling."""
import numpy as np
import pytest

import oracle
import pdcases as pc
from probabilistic_peridynamics_amd import _lib, autotune, batch, configs

pytestmark = pytest.mark.gpu

PARITY_RTOL = 1e-10


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    try:
        ok = _lib.device_count() > 0
    except OSError:
        ok = False
    if not ok:
        pytest.fail("no HIP device: the gpu tests need an MI355X")


def _samples(dt=None, shapes=((15, 3, 3), (20, 4, 4), (25, 5, 5),
                                (30, 6, 6))):
    ws = [configs.uq_sample(shape, s) for s, shape in enumerate(shapes)]
    if dt is not None:
        for w in ws:
            w.dt = dt
    return ws


def _oracle_run(w, steps):
    ref = oracle.SlotStepper(0, w.dt, w.coords, w.volume, w.family,
                             w.bc_types, w.bc_values, w.force_bc_types,
                             w.force_bc_values, w.nlist, w.bond_stiffness,
                             w.critical_stretch, cnode=w.extra["cnode"])
    d, f = w.magnitudes(steps)
    for s in range(steps):
        ref(d[s], f[s])
    return ref


@pytest.mark.parametrize("dt,steps", [(1e-5, 6), (5e-15, 40)])
def test_batch_samples_match_oracle(dt, steps):
    """dt = 1e-5 is the configured step, far past the overdamped Euler
    stability limit (~1e-14 s at dx = 0.01): bonds break within a few steps.
    5e-15 is below the limit, so rounding differences are not amplified and
    the forces are compared over a smooth run."""
    ws = _samples(dt)
    b = batch.SampleBatch(ws)
    b.run(steps, graph=True)
    res = b.results()
    for w, r in zip(ws, res):
        ref = _oracle_run(w, steps)
        nn, dmg = ref.connectivity()
        np.testing.assert_array_equal(r["nlist"], ref.nlist)
        np.testing.assert_array_equal(r["n_neigh"], nn)
        np.testing.assert_array_equal(r["damage"], dmg)
        assert pc.rel_l2(r["u"], ref.u) < PARITY_RTOL
        assert np.abs(ref.u).max() > 0


def test_batch_equals_solo_runs_bitwise():
    ws = _samples(1e-5)
    b = batch.SampleBatch(ws, lanes_per_node=4, tile_nodes=128)
    b.run(8, graph=True)
    res = b.results()
    for w, r in zip(ws, res):
        solo = batch.solo_integrator(w, lanes_per_node=4, tile_nodes=128)
        d, f = w.magnitudes(8)
        solo.run(8, d, f)
        u, _, _, _, _, dmg, nl, nn = solo.write(None, None, None, None, None,
                                                None, None, None)
        np.testing.assert_array_equal(r["u"], u)
        np.testing.assert_array_equal(r["damage"], dmg)
        np.testing.assert_array_equal(r["nlist"], nl)


def test_graph_runs_match_plain_launches_on_every_replay():
    ws = _samples(5e-15)
    g = batch.SampleBatch(ws)
    p = batch.SampleBatch(ws)
    for rep in range(3):
        g.run(5, graph=True)
        p.run(5, graph=False)
        p.integrator.synchronize()
        rg, rp = g.results(), p.results()
        for a, c in zip(rg, rp):
            np.testing.assert_array_equal(a["u"], c["u"])
            np.testing.assert_array_equal(a["nlist"], c["nlist"])


def test_tiles_never_span_samples_and_shapes_agree():
    ws = _samples(5e-15)
    ref = None
    for lp, t in [(4, 128), (2, 64), (8, 128), (1, 256)]:
        b = batch.SampleBatch(ws, lanes_per_node=lp, tile_nodes=t)
        stats = b.integrator.tile_stats()
        # every sample starts a new tile: at least one tile per sample
        assert stats["ntiles"] >= len(ws)
        b.run(10, graph=True)
        u = np.concatenate([r["u"] for r in b.results()])
        if ref is None:
            ref = u
        else:
            assert pc.rel_l2(u, ref) < 1e-13


def test_autotune_restores_initial_state():
    ws = _samples(5e-15)
    b = batch.SampleBatch(ws)
    best, table = autotune.autotune_batch(
        b, steps=4, cands=[(4, 128), (2, 256), (8, 64)])
    assert best in [c for c, _ in table]
    assert all(ms > 0 for _, ms in table)
    assert b.steps_done == 0
    b.run(6, graph=True)
    fresh = batch.SampleBatch(ws, lanes_per_node=best[0], tile_nodes=best[1])
    fresh.run(6, graph=True)
    for a, c in zip(b.results(), fresh.results()):
        np.testing.assert_array_equal(a["u"], c["u"])


def test_full_batch_of_256_samples_matches_the_oracle():
    """The whole batched workload (256 samples, 5,633,380 nodes) for six
    steps at its configured dt, one graph, every sample against its own
    oracle run: the broken set, lists and damage exactly, u to 1e-10."""
    ws = configs.uq_samples()
    assert len(ws) == 256 and sum(w.nnodes for w in ws) == 5633380
    b = batch.SampleBatch(ws)
    b.run(6, graph=True)
    worst = 0.0
    broke = 0
    for w, r in zip(ws, b.results()):
        ref = _oracle_run(w, 6)
        nn, dmg = ref.connectivity()
        np.testing.assert_array_equal(r["nlist"], ref.nlist)
        np.testing.assert_array_equal(r["n_neigh"], nn)
        np.testing.assert_array_equal(r["damage"], dmg)
        worst = max(worst, pc.rel_l2(r["u"], ref.u))
        broke += int(w.n_neigh.sum() - nn.sum())
    print(f"256 samples: worst u rel-L2 {worst:.2e}, {broke} bonds broken")
    assert worst < PARITY_RTOL and broke > 0
