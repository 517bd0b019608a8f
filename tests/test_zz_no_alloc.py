"""pd_monitor_noalloc (the monitor's sums with its scratch in a file-scope
device array, no allocation) against pd_monitor, through the C ABI: two
simulations in a row, the second smaller (fewer nodes, fewer tip groups),
so stale contents of the shared scratch would show."""
import numpy as np
import pytest

from probabilistic_peridynamics_amd import _lib, configs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    try:
        ok = _lib.device_count() > 0
    except OSError:
        ok = False
    if not ok:
        pytest.fail("no HIP device: the gpu tests need an MI355X")


def _sums(integ, fn, ngroups):
    model = np.full(6, np.nan)
    tips = np.full(5 * max(ngroups, 1), np.nan)
    _lib.check(fn(integ._sim, _lib.dptr(model), _lib.dptr(tips)))
    return model, tips[:5 * ngroups]


def _run(shape, ngroups, steps, seed):
    w = configs.uq_sample(shape, seed)
    w.dt = 5e-15
    w.disp_rate = 1e-5
    integ = configs.make_integrator(w)
    rng = np.random.default_rng(seed)
    n = w.nnodes
    node = rng.integers(0, n, 7 * ngroups).astype(np.int32)
    dof = rng.integers(0, 3, 7 * ngroups).astype(np.int32)
    group = np.repeat(np.arange(ngroups), 7).astype(np.int32)
    _lib.check(_lib.lib().pd_set_tips(integ._sim, len(node),
                                      _lib.iptr(node), _lib.iptr(dof),
                                      _lib.iptr(group), ngroups))
    d, f = w.magnitudes(steps)
    integ.run(steps, d, f)
    return integ


@pytest.mark.parametrize("order", ["large_then_small", "small_then_large"])
def test_noalloc_monitor_equals_pd_monitor(order):
    cases = [((40, 12, 12), 37, 12, 1), ((16, 5, 5), 3, 8, 2)]
    if order == "small_then_large":
        cases = cases[::-1]
    for shape, ngroups, steps, seed in cases:
        integ = _run(shape, ngroups, steps, seed)
        L = _lib.lib()
        want_m, want_t = _sums(integ, L.pd_monitor, ngroups)
        got_m, got_t = _sums(integ, L.pd_monitor_noalloc, ngroups)
        np.testing.assert_array_equal(got_m, want_m)
        np.testing.assert_array_equal(got_t, want_t)
        assert np.all(np.isfinite(got_m)) and np.any(got_m[:5] != 0.0)
        # and again, unchanged, from the array's last contents
        again_m, again_t = _sums(integ, L.pd_monitor_noalloc, ngroups)
        np.testing.assert_array_equal(again_m, want_m)
        np.testing.assert_array_equal(again_t, want_t)


def test_noalloc_monitor_rejects_too_many_tip_groups():
    w = configs.uq_sample((16, 5, 5), 3)
    integ = configs.make_integrator(w)
    g = 1025
    node = np.zeros(g, np.int32)
    dof = np.zeros(g, np.int32)
    group = np.arange(g, dtype=np.int32)
    L = _lib.lib()
    _lib.check(L.pd_set_tips(integ._sim, g, _lib.iptr(node),
                             _lib.iptr(dof), _lib.iptr(group), g))
    model, tips = np.zeros(6), np.zeros(5 * g)
    with pytest.raises(_lib.PDError, match="1024 tip groups"):
        _lib.check(L.pd_monitor_noalloc(integ._sim, _lib.dptr(model),
                                        _lib.dptr(tips)))
    # pd_monitor still serves it
    _lib.check(L.pd_monitor(integ._sim, _lib.dptr(model), _lib.dptr(tips)))
