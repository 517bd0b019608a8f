"""Built-in set-up predicates (predicates.PlaneBondTypes, predicates.Boxes)
against the reference's per-bond and per-node callback loops
(model.py:657-780, 1093-1212): the arrays a Model builds from them in bulk
equal the ones it builds calling them once per bond or node.  Host path
here (device -1); tests/test_gpu_predicates.py runs the device path."""
import warnings

import numpy as np
import pytest

import pdcases as pc
from probabilistic_peridynamics_amd import configs, predicates
from probabilistic_peridynamics_amd.model import Model

INF = np.inf


def composite_predicates(device):
    """Plane-split bond types (y = 0.5) and box boundary conditions, tips
    and densities on the reference's example mesh (1 x 1 plate)."""
    bt = predicates.PlaneBondTypes(1, 0.5, below=0, above=1, across=2,
                                   device=device)
    disp = predicates.Boxes([((-INF, -INF, -INF), (0.15, INF, INF),
                              [-1.0, 0.0, 0.0]),
                             ((0.85, -INF, -INF), (INF, INF, INF),
                              [1.0, 0.0, None])], device=device)
    force = predicates.Boxes([((0.45, 0.9, -INF), (0.55, INF, INF),
                               [None, -1e3, None])], device=device)
    tips = predicates.Boxes([((0.9, 0.4, -INF), (INF, 0.6, INF),
                              ["right", None, ("right", "mid")]),
                             ((0.4, 0.4, -INF), (0.6, 0.6, INF),
                              [None, "mid", None])], device=device)
    dens = predicates.Boxes([((-INF, 0.5, -INF), (INF, INF, INF), 2500.0)],
                            default=2400.0, device=device)
    return bt, disp, force, tips, dens


def _plain(pred):
    """The same predicate as an ordinary Python function (the reference's
    per-call form)."""
    def fn(*x):
        return pred(*x)
    return fn


def models(device, integrator):
    """(bulk, per-call) models of the example mesh with the predicates;
    ``integrator()`` makes each model's integrator."""
    bt, disp, force, tips, dens = composite_predicates(device)
    kw = dict(horizon=pc.HORIZON, critical_stretch=[[0.005, 0.008]] * 3,
              bond_stiffness=[[pc.BOND_STIFFNESS, 0.1 * pc.BOND_STIFFNESS]]
              * 3, dimensions=2)
    out = []
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        for wrap in (lambda p: p, _plain):
            out.append(Model(
                pc.GOLDEN / "example_mesh.vtk", integrator(),
                is_bond_type=wrap(bt), is_displacement_boundary=wrap(disp),
                is_force_boundary=wrap(force), is_tip=wrap(tips),
                is_density=wrap(dens), **kw))
    return out


def test_plane_bond_types_equal_the_per_bond_loop():
    coords = configs.lattice((12, 9, 7), 0.01, jitter=0.25, seed=3)
    fam, nl, nn = configs.neighbours(coords, 0.0301, "slot")
    p = predicates.PlaneBondTypes(1, 0.045, below=3, above=5, across=7,
                                  device=-1)
    got = p.bond_types(coords, nl, nn)
    want = np.zeros_like(got)
    for i in range(len(nn)):
        for s in range(nn[i]):
            want[i, s] = p(coords[i], coords[nl[i, s]])
    np.testing.assert_array_equal(got, want)
    assert set(np.unique(want[nl >= 0])) == {3, 5, 7}


def test_boxes_equal_the_per_node_calls():
    coords = configs.lattice((10, 10, 4), 0.1, jitter=0.3, seed=1)
    bt, disp, force, tips, dens = composite_predicates(-1)
    for pred in (disp, force, tips, dens):
        bulk = pred(coords.T.copy())
        for i, x in enumerate(coords):
            one = pred(x)
            if pred.scalar:
                assert bulk[i] == one
            else:
                for d in range(3):
                    if pred.numeric:
                        assert (np.isnan(bulk[d, i]) if one[d] is None
                                else bulk[d, i] == one[d])
                    else:
                        assert bulk[d, i] == one[d]


def _same_model(a, b):
    for name in ("bond_types", "bc_types", "bc_values", "force_bc_types",
                 "force_bc_values", "densities"):
        np.testing.assert_array_equal(getattr(a, name), getattr(b, name),
                                      err_msg=name)
    assert a.tip_types == b.tip_types
    assert a.ntips == b.ntips


def test_model_from_predicates_equals_per_call_model():
    """Model set-up with the predicates (bulk, host) and with the same
    predicates wrapped as plain functions (one call per bond / node)."""
    from test_host import Recorder
    bulk, plain = models(-1, lambda: Recorder(1e-3, "cl"))
    _same_model(bulk, plain)
    assert np.any(bulk.bond_types == 2) and np.any(bulk.force_bc_types)
    assert "right" in bulk.tip_types and "mid" in bulk.tip_types
