"""The two 1,038,230-node workloads against the oracle itself, step for
step: the beam (Velocity-Verlet, scalar stiffness, displacement supports
and a body-force strip) and the composite bar (three bond types x three
regimes, Euler-Cromer).  The oracle's per-node gather runs on the host's
cores (OpenMP), so ten full steps take seconds."""
import numpy as np
import pytest

import oracle
import pdcases as pc
from probabilistic_peridynamics_amd import _lib, configs

pytestmark = pytest.mark.gpu

KIND = {"euler": 0, "euler_cromer": 1, "velocity_verlet": 2}


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    try:
        ok = _lib.device_count() > 0
    except OSError:
        ok = False
    if not ok:
        pytest.fail("no HIP device: the gpu tests need an MI355X")


@pytest.mark.parametrize("name", ["beam_vv", "composite_ec"])
def test_full_size_against_the_oracle(name):
    w = getattr(configs, name)()
    integ = configs.make_integrator(w)
    ref = oracle.SlotStepper(
        KIND[w.integrator], w.dt, w.coords, w.volume, w.family, w.bc_types,
        w.bc_values, w.force_bc_types, w.force_bc_values, w.nlist,
        w.bond_stiffness, w.critical_stretch, damping=w.damping,
        densities=w.densities, bond_types=w.bond_types, plus_cs=w.plus_cs,
        nregimes=w.nregimes)
    d, f = w.magnitudes(10)
    integ.run(10, d, f)
    for s in range(10):
        ref(d[s], f[s])
    u, ud, udd, force, _, dmg, nl, nn = integ.write(
        None, None, None, None, None, None, None, None)
    rn, rd = ref.connectivity()
    np.testing.assert_array_equal(nl, ref.nlist)
    np.testing.assert_array_equal(nn, rn)
    np.testing.assert_array_equal(dmg, rd)
    if w.nregimes > 1:
        np.testing.assert_array_equal(integ.regimes(), ref.regimes)
    errs = {k: pc.rel_l2(a, b) for k, a, b in (
        ("u", u, ref.u), ("ud", ud, ref.ud), ("udd", udd, ref.udd),
        ("force", force, ref.force))}
    print(name, {k: f"{v:.2e}" for k, v in errs.items()})
    assert errs["u"] < 1e-10 and errs["ud"] < 1e-10
    assert errs["udd"] < 1e-10 and errs["force"] < 1e-10
    assert np.any(u != 0.0)


@pytest.mark.parametrize("name", ["beam_vv", "composite_ec"])
def test_full_size_breaking_against_the_oracle(name):
    """The same workloads loaded hard enough that, within ten steps, bonds
    break across the 1M-node mesh (the beam's body force scaled by 5e9)
    and walk their regimes (the composite's ends pulled 2e-5 m per step):
    lists, damage and regimes exact, u to 1e-10."""
    w = getattr(configs, name)()
    if name == "beam_vv":
        w.force_scale = 5e9
    else:
        w.disp_rate = 2e-5
    integ = configs.make_integrator(w)
    ref = oracle.SlotStepper(
        KIND[w.integrator], w.dt, w.coords, w.volume, w.family, w.bc_types,
        w.bc_values, w.force_bc_types, w.force_bc_values, w.nlist,
        w.bond_stiffness, w.critical_stretch, damping=w.damping,
        densities=w.densities, bond_types=w.bond_types, plus_cs=w.plus_cs,
        nregimes=w.nregimes)
    d, f = w.magnitudes(10)
    integ.run(10, d, f)
    for s in range(10):
        ref(d[s], f[s])
    u, ud, udd, force, _, dmg, nl, nn = integ.write(
        None, None, None, None, None, None, None, None)
    rn, rd = ref.connectivity()
    np.testing.assert_array_equal(nl, ref.nlist)
    np.testing.assert_array_equal(nn, rn)
    np.testing.assert_array_equal(dmg, rd)
    broken = w.live_bonds() - int(rn.sum())
    print(name, "broken", broken)
    assert broken > 1000
    if w.nregimes > 1:
        np.testing.assert_array_equal(integ.regimes(), ref.regimes)
        walked = int(np.count_nonzero(ref.regimes[w.nlist >= 0]))
        print(name, "bonds past regime 0", walked)
        assert walked > 1000
    errs = {k: pc.rel_l2(a, b) for k, a, b in (
        ("u", u, ref.u), ("ud", ud, ref.ud), ("force", force, ref.force))}
    print(name, {k: f"{v:.2e}" for k, v in errs.items()})
    assert errs["u"] < 1e-10 and errs["ud"] < 1e-10
    assert errs["force"] < 1e-9
