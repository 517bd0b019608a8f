"""Batched UQ samples on the host: sizes, stiffness fields, concatenation,
autotuner logic (no GPU)."""
import numpy as np
import pytest

from probabilistic_peridynamics_amd import autotune, batch, configs


def test_uq_sizes_are_log_uniform_bars():
    sizes = configs.uq_sizes()
    n = np.array([a * b * c for a, b, c in sizes])
    assert len(sizes) == 256
    assert all(a == 5 * b and b == c for a, b, c in sizes)
    assert n.min() >= 5000 and n.max() <= 61000
    # log-uniform on [5e3, 6e4]: mean (6e4 - 5e3) / ln 12 ~ 22.1k
    assert 19000 < n.mean() < 25000
    assert configs.uq_sizes() == sizes   # seeded


def test_gaussian_random_field_statistics():
    rng = np.random.default_rng(7)
    dx, ell, sigma = 0.01, 0.05, 0.3
    z = np.stack([configs.gaussian_random_field((48, 16, 16), dx, ell, sigma,
                                                rng) for _ in range(60)])
    assert z.std() == pytest.approx(sigma, rel=0.08)
    assert abs(z.mean()) < 0.05
    # squared-exponential correlation at lags of 1, 5 and 10 cells
    for lag in (1, 5, 10):
        rho = np.mean(z[:, :-lag] * z[:, lag:]) / z.var()
        assert rho == pytest.approx(np.exp(-(lag * dx) ** 2 / (2 * ell ** 2)),
                                    abs=0.08)


def _tiny_samples():
    out = []
    for s, shape in enumerate([(10, 2, 2), (15, 3, 3), (5, 1, 1)]):
        out.append(configs.uq_sample(shape, s))
    return out


def test_concatenate_offsets_and_groups():
    ws = _tiny_samples()
    cat = batch.concatenate(ws)
    n = sum(w.nnodes for w in ws)
    assert cat["coords"].shape == (n, 3)
    assert list(cat["offsets"]) == [0, 40, 175, 180]
    assert np.array_equal(np.bincount(cat["group"]), [40, 135, 5])
    nl = cat["nlist"]
    assert nl.shape[1] == max(w.max_neighbours for w in ws)
    # no bond joins two samples; rows are the samples' rows, offset
    for w, o, g in zip(ws, cat["offsets"][:-1], range(len(ws))):
        rows = nl[o:o + w.nnodes]
        live = rows[rows >= 0]
        assert np.all(cat["group"][live] == g)
        assert np.array_equal(
            np.where(rows[:, :w.max_neighbours] >= 0,
                     rows[:, :w.max_neighbours] - o, -1), w.nlist)
    assert cat["n_neigh"].sum() == sum(w.n_neigh.sum() for w in ws)


def test_uq_sample_field_and_loading():
    w = configs.uq_sample((20, 4, 4), 3)
    c = configs.micromodulus(w.horizon)
    assert np.allclose(w.extra["cnode"], c * np.exp(w.extra["g"]))
    x = w.coords[:, 0]
    assert np.all(w.bc_types[x < 0.02] == 1)
    pulled = x > 0.2 - 0.02
    assert np.all(w.bc_values[pulled, 0] == 1.0)
    assert w.dt == 1e-5 and w.critical_stretch == 0.005


def test_autotune_picks_fastest_and_applies_it():
    applied = []
    times = {(1, 128): 5.0, (2, 128): 3.0, (4, 64): 4.0}
    state = {}

    def apply(lp, t):
        applied.append((lp, t))
        state["cur"] = (lp, t)

    best, table = autotune.autotune(apply, lambda: times[state["cur"]],
                                    cands=list(times))
    assert best == (2, 128)
    assert applied[-1] == (2, 128)
    assert dict(table) == times


def test_autotune_candidates_respect_workgroup_limit():
    for lp, t in autotune.candidates():
        assert 128 <= lp * t <= 1024 and t % 64 == 0 and lp in (1, 2, 4, 8)
