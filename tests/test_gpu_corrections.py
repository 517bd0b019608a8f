"""Stiffness corrections on the GPU (pd_set_* with device 0) against the
oracle's restatement of peripy/correction.pyx: bitwise, on lists with
repeats and asymmetry, on the reference's example meshes, and on the full
1,038,230-node beam."""
import time
import warnings

import numpy as np
import pytest

import oracle
from probabilistic_peridynamics_amd import _lib, configs
from probabilistic_peridynamics_amd import correction as corr
from probabilistic_peridynamics_amd import integrators as it
from probabilistic_peridynamics_amd.model import Model

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    try:
        ok = _lib.device_count() > 0
    except OSError:
        ok = False
    if not ok:
        pytest.fail("no HIP device: the gpu tests need an MI355X")


def _all_four(coords, nl, nn, vol, horizon, radius, bulk, base):
    """[(name, oracle result, device result)] for the four functions."""
    cases = [
        ("set_micromodulus_function", (coords, nl, nn, horizon, 0)),
        ("set_volume_correction", (coords, nl, nn, horizon, radius, 0)),
        ("set_imprecise_surface_correction",
         (nl, nn, np.float64(np.sum(vol) / len(vol)), bulk)),
        ("set_precise_surface_correction", (nl, nn, vol, bulk)),
    ]
    out = []
    for fn, args in cases:
        a, b = base.copy(), base.copy()
        getattr(oracle, fn)(a, *args)
        getattr(corr, fn)(b, *args, device=0)
        out.append((fn, a, b))
    return out


def test_random_lists_bitwise():
    """Repeats, asymmetric rows, self bonds and out-of-range entries."""
    rng = np.random.default_rng(11)
    for trial in range(40):
        n = int(rng.integers(1, 300))
        k = int(rng.integers(1, 40))
        nl = rng.integers(-2, n + 2, size=(n, k)).astype(np.intc)
        nn = rng.integers(0, k + 1, size=n).astype(np.intc)
        coords = rng.random((n, 3))
        vol = rng.random(n)
        base = rng.random((n, k)) + 0.5
        for fn, a, b in _all_four(coords, nl, nn, vol, rng.random(),
                                  0.2 * rng.random(), 2.5, base):
            assert np.array_equal(a, b), (fn, trial, n, k)


@pytest.mark.parametrize("layout", ["cython", "cl"])
@pytest.mark.parametrize("dim,mesh", [(2, "example_mesh.vtk"),
                                      (3, "example_mesh_3d.vtk")])
def test_example_meshes_bitwise(golden, dim, mesh, layout):
    from test_host import Recorder
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        m = Model(golden / mesh, Recorder(1e-3, layout), horizon=0.1,
                  critical_stretch=0.05, bond_stiffness=1.0, dimensions=dim)
    nl, nn = m.initial_connectivity
    nl = np.ascontiguousarray(nl, dtype=np.intc)
    nn = np.ascontiguousarray(nn, dtype=np.intc)
    bulk = np.pi * 0.01 if dim == 2 else (4. / 3) * np.pi * 1e-3
    for fn, a, b in _all_four(m.coords, nl, nn, m.volume, m.horizon,
                              np.power(1. / m.nnodes, 1. / 3), bulk,
                              np.ones(nl.shape)):
        assert np.array_equal(a, b), fn


def test_model_with_hip_integrator_uses_the_gpu(golden):
    """Model with a HIP integrator computes the superimposed corrections on
    its GPU: bitwise those of the host path, in the reference's order."""
    kw = dict(horizon=0.1, critical_stretch=0.05, bond_stiffness=1.0,
              surface_correction=1, micromodulus_function=0,
              volume_correction=0, node_radius=0.01)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        m = Model(golden / "example_mesh.vtk", it.EulerHIP(1e-3), **kw)
    assert m._setup_device() == 0
    nl, nn = m.initial_connectivity
    b = np.ones(nl.shape)
    corr.set_micromodulus_function(b, m.coords, nl, nn, 0.1, 0, device=-1)
    corr.set_volume_correction(b, m.coords, nl, nn, 0.01, 0.1, 0, device=-1)
    corr.set_precise_surface_correction(
        b, nl, nn, m.volume, np.float64(np.pi * np.power(m.horizon, 2)),
        device=-1)
    assert np.array_equal(m.stiffness_corrections, b)


def test_full_size_beam_bitwise():
    """The 470 x 47 x 47 beam (1,038,230 nodes, 109,555,662 bonds): the
    chain micromodulus -> volume -> precise surface, and the imprecise
    surface factors, bitwise against the oracle."""
    w = configs.beam_vv()
    nl = np.ascontiguousarray(w.nlist, dtype=np.intc)
    nn = np.ascontiguousarray(w.n_neigh, dtype=np.intc)
    bulk = np.float64((4. / 3) * np.pi * np.power(w.horizon, 3))
    radius = 0.3 * w.horizon
    a = np.ones(nl.shape)
    t0 = time.perf_counter()
    oracle.set_micromodulus_function(a, w.coords, nl, nn, w.horizon, 0)
    oracle.set_volume_correction(a, w.coords, nl, nn, w.horizon, radius, 0)
    oracle.set_precise_surface_correction(a, nl, nn, w.volume, bulk)
    t_cpu = time.perf_counter() - t0
    b = np.ones(nl.shape)
    t0 = time.perf_counter()
    corr.set_micromodulus_function(b, w.coords, nl, nn, w.horizon, 0,
                                   device=0)
    corr.set_volume_correction(b, w.coords, nl, nn, w.horizon, radius, 0,
                               device=0)
    corr.set_precise_surface_correction(b, nl, nn, w.volume, bulk, device=0)
    t_gpu = time.perf_counter() - t0
    print(f"corrections chain: oracle {t_cpu:.2f} s, GPU {t_gpu:.2f} s "
          f"(with transfers) for {w.nnodes} nodes")
    assert np.array_equal(a, b)
    a = np.ones(nl.shape)
    b = np.ones(nl.shape)
    avg = np.float64(np.sum(w.volume) / w.nnodes)
    oracle.set_imprecise_surface_correction(a, nl, nn, avg, bulk)
    corr.set_imprecise_surface_correction(b, nl, nn, avg, bulk, device=0)
    assert np.array_equal(a, b)
