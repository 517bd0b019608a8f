"""The set-up predicates evaluated on the GPU (pd_plane_bond_types,
pd_first_box) against the reference's per-bond and per-node calls: bond
types, boundary conditions, tips and densities a Model builds from them
equal those it builds calling them one bond or node at a time, and the
bulk arrays equal the per-call ones on a 100k-node jittered bar."""
import numpy as np
import pytest

from probabilistic_peridynamics_amd import _lib, configs, predicates
from test_predicates import composite_predicates, models, _same_model

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    try:
        ok = _lib.device_count() > 0
    except OSError:
        ok = False
    if not ok:
        pytest.fail("no HIP device: the gpu tests need an MI355X")


def test_device_bond_types_equal_host_and_per_bond_calls():
    coords = configs.lattice((100, 32, 32), 0.01, jitter=0.25, seed=5,
                             origin=(0.0, -0.16, 0.0))
    fam, nl, nn = configs.neighbours(coords, 0.0301, "slot")
    dev = predicates.PlaneBondTypes(1, 0.0, below=0, above=1, across=2,
                                    device=0)
    host = predicates.PlaneBondTypes(1, 0.0, below=0, above=1, across=2,
                                     device=-1)
    got = dev.bond_types(coords, nl, nn)
    np.testing.assert_array_equal(got, host.bond_types(coords, nl, nn))
    # per-bond calls on a sample of rows
    rows = np.random.default_rng(1).choice(len(nn), 2000, replace=False)
    for i in rows:
        for s in range(nn[i]):
            assert got[i, s] == dev(coords[i], coords[nl[i, s]])
        assert np.all(got[i, nn[i]:] == 0)
    assert set(np.unique(got[nl >= 0])) == {0, 1, 2}


def test_device_boxes_equal_per_node_calls():
    coords = configs.lattice((60, 30, 20), 0.02, jitter=0.3, seed=2)
    x = np.ascontiguousarray(coords * np.array([1 / 1.2, 1 / 0.6, 1 / 0.4]))
    for pred in composite_predicates(0)[1:]:
        bulk = pred(x.T.copy())
        for i in np.random.default_rng(3).choice(len(x), 3000,
                                                 replace=False):
            one = pred(x[i])
            if pred.scalar:
                assert bulk[i] == one
            elif pred.numeric:
                for d in range(3):
                    assert (np.isnan(bulk[d, i]) if one[d] is None
                            else bulk[d, i] == one[d])
            else:
                assert list(bulk[:, i]) == one


def test_model_with_device_predicates_equals_per_call_model():
    """A Model on the GPU: bond types, BCs, tips and densities from the
    device predicates equal those of the same predicates called per bond
    and per node."""
    from probabilistic_peridynamics_amd.integrators import EulerCromerHIP
    bulk, plain = models(0, lambda: EulerCromerHIP(0.0, 1e-3))
    _same_model(bulk, plain)
    assert np.any(bulk.bond_types == 2) and np.any(bulk.force_bc_types)
