"""Periodic writes on the device (pd_set_tips / pd_monitor, used by
Model.simulate with a HIP integrator) against the reference's host loop of
model.py:1303-1360 run on full copies of the same state."""
import warnings

import numpy as np
import pytest

import pdcases as pc
from probabilistic_peridynamics_amd import _lib
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


def is_tip(x):
    """Three tip groups; the x-DOF of right-hand nodes is in two."""
    tip = [None, None, None]
    if x[0] > 0.9:
        tip[0] = ("right", "edge")
        tip[1] = "right"
    elif x[1] > 0.95:
        tip[2] = "top"
        tip[0] = "edge"
    return tip


def is_force_boundary(x):
    return [None, 1.0e-3, None] if x[0] > 0.95 else [None, None, None]


def _model(golden):
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        return Model(golden / "example_mesh.vtk",
                     it.VelocityVerletHIP(damping=1.0, dt=1e-3),
                     horizon=pc.HORIZON, critical_stretch=0.005,
                     bond_stiffness=pc.BOND_STIFFNESS,
                     is_displacement_boundary=pc.simple_displacement_boundary,
                     is_force_boundary=is_force_boundary, is_tip=is_tip,
                     density=np.ones(2113))


def test_simulate_writes_match_the_host_loop(golden, tmp_path):
    """The same run written by the device sums and by the reference's host
    loop over full copies: tip sums bitwise, model sums to rounding, the
    mesh files and the final state identical."""
    mags = 1e-5 * np.arange(1, 31)
    (tmp_path / "dev").mkdir()
    (tmp_path / "host").mkdir()
    m_dev = _model(golden)
    out_dev = m_dev.simulate(30, displacement_bc_magnitudes=mags,
                             force_bc_magnitudes=np.ones(30), write=5,
                             write_path=tmp_path / "dev")
    m_host = _model(golden)
    m_host.integrator.monitor = None  # forces the full-write loop
    out_host = m_host.simulate(30, displacement_bc_magnitudes=mags,
                               force_bc_magnitudes=np.ones(30), write=5,
                               write_path=tmp_path / "host")
    data_dev, data_host = out_dev[5], out_host[5]
    assert set(data_dev) == set(data_host) == {"model", "right", "edge",
                                               "top"}
    for tip in ("right", "edge", "top"):
        for k, v in data_host[tip].items():
            assert np.array_equal(data_dev[tip][k], v), (tip, k)
    assert np.array_equal(data_dev["model"]["step"],
                          data_host["model"]["step"])
    # the model sums cancel heavily (|sum| << sum of |terms|), so they are
    # compared against the size of their terms, taken from the final state
    # (acceleration sums are checked term-scaled in the next test)
    u, ud, force = out_host[0], out_host[4], out_host[3]
    v = m_host.volume[:, None]
    scale = {"displacement": np.sum(np.abs(u)) / m_host.nnodes,
             "velocity": np.sum(np.abs(ud)) / m_host.nnodes,
             "force": np.sum(np.abs(force * v)),
             "body_force": np.sum(np.abs(force * v))}
    for k, sc in scale.items():
        np.testing.assert_allclose(data_dev["model"][k],
                                   data_host["model"][k], rtol=0,
                                   atol=1e-13 * sc, err_msg=k)
    np.testing.assert_allclose(data_dev["model"]["damage_sum"],
                               data_host["model"]["damage_sum"], rtol=1e-12)
    for a, b in zip(out_dev[:2] + out_dev[3:5], out_host[:2] + out_host[3:5]):
        assert np.array_equal(a, b)
    assert np.array_equal(out_dev[2][0], out_host[2][0])
    for step in range(5, 31, 5):
        f = f"U_{step}.vtk"
        assert ((tmp_path / "dev" / f).read_bytes()
                == (tmp_path / "host" / f).read_bytes())


def test_monitor_direct_sums(golden):
    """pd_monitor after a few steps against sums of a full pd_write, with
    interleaved groups and a DOF listed twice."""
    m = _model(golden)
    integ = m.integrator
    m.simulate(3, displacement_bc_magnitudes=1e-4 * np.ones(3))
    rng = np.random.default_rng(2)
    tips = {"a": [], "b": [], "c": []}
    for _ in range(300):
        g = "abc"[rng.integers(3)]
        tips[g].append((int(rng.integers(m.nnodes)), int(rng.integers(3))))
    tips["a"].append(tips["a"][0])
    integ.set_tips(tips)
    sums, tip_sums = integ.monitor()
    u, ud, udd, force, body, damage, _, _ = integ.write(
        None, None, None, None, None, None, None, None)
    v = m.volume
    for g, entries in tips.items():
        acc = np.zeros(5)
        for i, j in entries:
            acc += [u[i, j], ud[i, j], udd[i, j], force[i, j] * v[i],
                    body[i, j] * v[i]]
        # the loop above adds the five at once in the same order
        got = [tip_sums[g][k] for k in integ._SUMS]
        assert np.array_equal(got, acc), g
    ref = [np.sum(u), np.sum(ud), np.sum(udd), np.sum(force * v[:, None]),
           np.sum(body * v[:, None])]
    scale = [np.sum(np.abs(x)) for x in (u, ud, udd, force * v[:, None],
                                         body * v[:, None])]
    for k, r, sc in zip(integ._SUMS, ref, scale):
        assert abs(sums[k] - r) <= 1e-14 * sc + 1e-300, k
    assert abs(sums["damage_sum"] - np.sum(damage)) <= 1e-12


def test_cython_euler_keeps_the_body_force_it_was_given(golden, tmp_path):
    """integrators.py:371-380, 456-459: the Cython Euler never fills
    body_force, so its write returns the array it was given (zeros from
    simulate) and the written body-force sums are 0; ud and udd likewise."""
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        m = Model(golden / "example_mesh.vtk", it.Euler(1e-3),
                  horizon=pc.HORIZON, critical_stretch=0.005,
                  bond_stiffness=pc.BOND_STIFFNESS,
                  is_displacement_boundary=pc.simple_displacement_boundary,
                  is_force_boundary=is_force_boundary, is_tip=is_tip)
    (tmp_path / "w").mkdir()
    out = m.simulate(10, displacement_bc_magnitudes=1e-5 * np.arange(1, 11),
                     force_bc_magnitudes=np.ones(10), write=5,
                     write_path=tmp_path / "w")
    data = out[5]
    assert np.all(data["model"]["body_force"] == 0.0)
    for tip in ("right", "edge", "top"):
        assert np.all(data[tip]["body_force"] == 0.0)
        assert np.all(data[tip]["velocity"] == 0.0)
    assert np.any(data["model"]["force"] != 0.0)
    st = m.integrator.write(None, None, None, None, None, None, None, None)
    assert np.all(st[4] == 0.0) and np.all(st[1] == 0.0)
    assert np.any(st[3] != 0.0)


def test_simulate_in_library_loops_equals_per_step_calls(golden, tmp_path):
    """Model.simulate hands the steps between two writes to the library's
    step loop; the result equals one Python call per step bitwise,
    whatever the write period, first step and run length."""
    mags = 1e-5 * np.arange(1, 40)
    outs = []
    for per_step in (False, True):
        m = _model(golden)
        if per_step:
            m.integrator.run = None
        path = tmp_path / ("step" if per_step else "loop")
        path.mkdir()
        out = m.simulate(23, displacement_bc_magnitudes=mags,
                         force_bc_magnitudes=np.ones(39), write=7,
                         first_step=4, write_path=path)
        outs.append((out, sorted(p.name for p in path.iterdir())))
    (a, fa), (b, fb) = outs
    assert fa == fb == ["U_14.vtk", "U_21.vtk", "U_7.vtk"]
    for x, y in zip(a[:2] + a[3:5], b[:2] + b[3:5]):
        assert np.array_equal(x, y)
    assert np.array_equal(a[2][0], b[2][0])
    for key in a[5]:
        for k, v in a[5][key].items():
            assert np.array_equal(v, b[5][key][k]), (key, k)


def test_partial_writes_equal_the_full_write(golden):
    """pd_write with some pointers NULL (body_force without force, which
    takes force where no force BC applies; ud alone; u and damage alone)
    gives what the full write gives, for the same state."""
    dptr = _lib.dptr
    m = _model(golden)
    m.simulate(4, displacement_bc_magnitudes=1e-4 * np.ones(4),
               force_bc_magnitudes=np.ones(4))
    integ = m.integrator
    full = integ.write(None, None, None, None, None, None, None, None)
    n = m.nnodes
    bf = np.full((n, 3), np.nan)
    ud = np.full((n, 3), np.nan)
    _lib.check(_lib.lib().pd_write(integ._sim, None, None, None, None,
                                   dptr(bf), None, None, None))
    _lib.check(_lib.lib().pd_write(integ._sim, None, dptr(ud), None, None,
                                   None, None, None, None))
    assert np.array_equal(bf, full[4]) and np.array_equal(ud, full[1])
    assert np.any(full[4] != full[3])  # force BCs make them differ
    u, dmg = integ.write_fields(None, None)
    assert np.array_equal(u, full[0]) and np.array_equal(dmg, full[5])
