"""The GPU adjoint (pd_adjoint_*) against the torch reference gradient and
central finite differences; checkpoint-interval independence; state kept."""
import numpy as np
import pytest

import adjoint_ref
import pdcases as pc
from probabilistic_peridynamics_amd import _lib, adjoint
from test_adjoint_ref import adjoint_workload

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    try:
        ok = _lib.device_count() > 0
    except OSError:
        ok = False
    if not ok:
        pytest.fail("no HIP device: the gpu tests need an MI355X")


def _setup(steps=40, every=16, shape=(16, 5, 5), seed=0):
    w = adjoint_workload(shape=shape, seed=seed)
    rng = np.random.default_rng(11)
    sensors = rng.choice(w.nnodes, 24, replace=False)
    target = rng.normal(scale=1e-6, size=(24, 3))
    solver = adjoint.StiffnessAdjoint(w, steps, checkpoint_every=every)
    return w, solver, sensors, target


def _loss_grad(solver, g, sensors, target):
    u = solver.forward(g)
    dL = np.zeros_like(u)
    dL[sensors] = 2.0 * (u[sensors] - target)
    return float(((u[sensors] - target) ** 2).sum()), solver.backward(dL), u


def test_adjoint_matches_reference_gradient():
    w, solver, sensors, target = _setup()
    g = w.extra["g"]
    L, grad, u = _loss_grad(solver, g, sensors, target)
    Lr, grad_r, ur, alive = adjoint_ref.loss_and_grad(
        w, g, w.bond_stiffness, solver.steps, sensors, target)
    # same trajectory and the same broken bonds
    _, _, nl = solver.state()
    ii, jj = adjoint_ref.edges(w.nlist)
    cols = np.nonzero(w.nlist >= 0)[1]
    np.testing.assert_array_equal(nl[ii, cols] >= 0, alive)
    assert (~alive).sum() > 0
    assert pc.rel_l2(u, ur) < 1e-12
    assert L == pytest.approx(Lr, rel=1e-10)
    assert pc.rel_l2(grad, grad_r) < 1e-10


def test_adjoint_matches_central_differences():
    w, solver, sensors, target = _setup()
    g = w.extra["g"]
    L, grad, _ = _loss_grad(solver, g, sensors, target)
    rng = np.random.default_rng(5)
    h = 1e-6
    worst = 0.0
    for _ in range(8):
        v = rng.standard_normal(w.nnodes)
        Lp = _loss_grad(solver, g + h * v, sensors, target)[0]
        Lm = _loss_grad(solver, g - h * v, sensors, target)[0]
        fd = (Lp - Lm) / (2 * h)
        worst = max(worst, abs(grad @ v - fd) / abs(fd))
    assert worst < 1e-6


def test_gradient_independent_of_checkpoint_interval():
    grads = []
    for every in (1, 7, 40, 100):
        w, solver, sensors, target = _setup(every=every)
        grads.append(_loss_grad(solver, w.extra["g"], sensors, target)[1])
    for gr in grads[1:]:
        np.testing.assert_array_equal(gr, grads[0])


def test_backward_keeps_the_final_state_and_repeats():
    w, solver, sensors, target = _setup()
    g = w.extra["g"]
    L, grad, u = _loss_grad(solver, g, sensors, target)
    u1, d1, nl1 = solver.state()
    dL = np.zeros_like(u)
    dL[sensors] = 2.0 * (u[sensors] - target)
    grad2 = solver.backward(dL)
    u2, d2, nl2 = solver.state()
    np.testing.assert_array_equal(grad, grad2)
    np.testing.assert_array_equal(u1, u)
    np.testing.assert_array_equal(u2, u)
    np.testing.assert_array_equal(nl2, nl1)
    np.testing.assert_array_equal(d2, d1)


def test_torch_autograd_function():
    torch = pytest.importorskip("torch")
    w, solver, sensors, target = _setup()
    g = torch.tensor(w.extra["g"], dtype=torch.float64, requires_grad=True)
    L = adjoint.sensor_loss(g, solver, torch.as_tensor(sensors),
                            torch.as_tensor(target))
    L.backward()
    Lr, grad_r, _, _ = adjoint_ref.loss_and_grad(
        w, w.extra["g"], w.bond_stiffness, solver.steps, sensors, target)
    assert float(L) == pytest.approx(Lr, rel=1e-10)
    assert pc.rel_l2(g.grad.numpy(), grad_r) < 1e-10


def test_gradient_at_a_large_step_with_nodes_at_rest():
    """dt = 1e-5 on the stiff bar: k dt ~ 1e16, nodes away from the pulled
    face stay exactly at rest while its neighbours fly off and break.  A
    bond with l == l0 bitwise must get 1/l0 - 1/l = 0 exactly (its forward
    stretch is 0): the gradient matches central differences."""
    w = adjoint_workload(shape=(30, 10, 10), dt=1e-5, pull=1e-8, seed=4)
    rng = np.random.default_rng(4)
    sensors = rng.choice(w.nnodes, 24, replace=False)
    solver = adjoint.StiffnessAdjoint(w, 30, checkpoint_every=8)
    g = w.extra["g"]
    target = solver.forward(0.5 * g)[sensors]
    L, grad, u = _loss_grad(solver, g, sensors, target)
    _, dmg, _ = solver.state()
    assert np.all(np.isfinite(grad)) and 0 < np.mean(dmg) < 0.5
    h = 1e-6
    worst = 0.0
    for _ in range(4):
        v = rng.standard_normal(w.nnodes)
        Lp = _loss_grad(solver, g + h * v, sensors, target)[0]
        Lm = _loss_grad(solver, g - h * v, sensors, target)[0]
        fd = (Lp - Lm) / (2 * h)
        worst = max(worst, abs(grad @ v - fd) / abs(fd))
    assert worst < 1e-6


def test_torch_autograd_on_device_tensors():
    """g, u and the gradient stay on the GPU (the library takes them by
    device address): the gradient equals the reference gradient and the
    host path's."""
    torch = pytest.importorskip("torch")
    w, solver, sensors, target = _setup()
    dev = torch.device("cuda", 0)
    g = torch.tensor(w.extra["g"], dtype=torch.float64, device=dev,
                     requires_grad=True)
    u = adjoint.displacements(g, solver)
    assert u.is_cuda
    L = ((u[torch.as_tensor(sensors, device=dev)]
          - torch.as_tensor(target, device=dev)) ** 2).sum()
    L.backward()
    assert g.grad.is_cuda
    Lr, grad_r, _, _ = adjoint_ref.loss_and_grad(
        w, w.extra["g"], w.bond_stiffness, solver.steps, sensors, target)
    assert float(L) == pytest.approx(Lr, rel=1e-10)
    assert pc.rel_l2(g.grad.cpu().numpy(), grad_r) < 1e-10
    _, grad_host, _ = _loss_grad(solver, w.extra["g"], sensors, target)
    assert pc.rel_l2(g.grad.cpu().numpy(), grad_host) < 1e-12


@pytest.mark.parametrize("device", ["cpu", "cuda"])
def test_two_evaluations_before_backward(device):
    """Two forwards, then the first loss's backward: the solver's recording
    belongs to the second field, so the first node re-records its own; its
    gradient equals the one taken alone."""
    torch = pytest.importorskip("torch")
    w, solver, sensors, target = _setup()
    dev = torch.device(device, 0) if device == "cuda" else torch.device("cpu")
    s_t = torch.as_tensor(sensors, device=dev)
    t_t = torch.as_tensor(target, device=dev)

    def loss(gg):
        return ((adjoint.displacements(gg, solver)[s_t] - t_t) ** 2).sum()

    g0 = np.asarray(w.extra["g"])
    alone = torch.tensor(g0, dtype=torch.float64, device=dev,
                         requires_grad=True)
    loss(alone).backward()
    g1 = torch.tensor(g0, dtype=torch.float64, device=dev,
                      requires_grad=True)
    g2 = torch.tensor(0.5 * g0, dtype=torch.float64, device=dev,
                      requires_grad=True)
    L1, L2 = loss(g1), loss(g2)
    L1.backward()
    np.testing.assert_array_equal(g1.grad.cpu().numpy(),
                                  alone.grad.cpu().numpy())
    L2.backward()
    assert pc.rel_l2(g2.grad.cpu().numpy(), alone.grad.cpu().numpy()) > 1e-3


def test_full_size_adjoint_against_central_differences():
    """The adjoint configuration at full size: the 50,820-node jittered bar
    (105 x 22 x 22), 2000 Euler steps at dt = 1e-5 with checkpoints every
    100, 64 sensors, targets from a perturbed field (GRF sigma 0.1); the
    gradient of the on-device autograd function against central
    differences (h = 1e-6) on four random directions."""
    torch = pytest.importorskip("torch")
    from probabilistic_peridynamics_amd import configs
    w = configs.uq_sample((105, 22, 22), 4)
    rng = np.random.default_rng(4)
    g_true = configs.gaussian_random_field((105, 22, 22), 0.01, 0.05, 0.1,
                                           rng).reshape(-1)
    sensors = rng.choice(w.nnodes, 64, replace=False)
    solver = adjoint.StiffnessAdjoint(w, 2000, checkpoint_every=100)
    target = solver.forward(g_true)[sensors]
    dev = torch.device("cuda", 0)
    s_t = torch.as_tensor(sensors, device=dev)
    t_t = torch.as_tensor(target, device=dev)
    g = torch.zeros(w.nnodes, dtype=torch.float64, device=dev,
                    requires_grad=True)
    L = ((adjoint.displacements(g, solver)[s_t] - t_t) ** 2).sum()
    L.backward()
    grad = g.grad.cpu().numpy()
    assert np.all(np.isfinite(grad))

    def loss(gg):
        u = solver.forward(gg)
        return float(((u[sensors] - target) ** 2).sum())

    h = 1e-6
    errs = []
    for _ in range(4):
        v = rng.standard_normal(w.nnodes)
        fd = (loss(h * v) - loss(-h * v)) / (2 * h)
        errs.append(abs(grad @ v - fd) / abs(fd))
    print(f"loss {float(L):.4e}, FD relative errors "
          + " ".join(f"{e:.2e}" for e in errs))
    assert max(errs) <= 1e-6
