"""Gradients of displacement losses with respect to the stiffness field.

The forward model is the Euler step of a :class:`~.configs.Workload` with a
per-node stiffness field ``c = c0 exp(g)`` (bond stiffness ``(c_i + c_j)/2``)
run on the GPU.  The backward pass is the library's hand-written adjoint
(``pd_adjoint_*``): the forward pass records the step at which each bond
broke and checkpoints ``u``; the backward pass replays each checkpoint
segment with the step kernel and runs the adjoint kernel back through it,
holding the recorded break mask fixed.  :class:`StiffnessField` wraps this
as a ``torch.autograd.Function`` of the log-stiffness ``g``.
"""
import numpy as np

from . import integrators as it


class StiffnessAdjoint:
    """Forward and adjoint solves of one workload for fields g."""

    def __init__(self, w, steps, checkpoint_every=100, c0=None,
                 context=None, lanes_per_node=4, tile_nodes=128,
                 tension_only=False):
        if w.integrator != "euler":
            raise ValueError("the adjoint is for the Euler integrator")
        self.w = w
        self.steps = int(steps)
        self.every = int(checkpoint_every)
        self.c0 = float(w.bond_stiffness if c0 is None else c0)
        self.integrator = it.EulerHIP(w.dt, context=context)
        self.integrator.build(w.nnodes, 3, w.max_neighbours, w.coords,
                              w.volume, w.family, w.bc_types, w.bc_values,
                              w.force_bc_types, w.force_bc_values, None,
                              None, None)
        self.integrator.set_break_mode(tension_only)
        self.integrator.set_launch(lanes_per_node, tile_nodes)
        self.dmags, self.fmags = w.magnitudes(self.steps)
        self.c = None

    def forward(self, g):
        """u after ``steps`` steps from rest, for c = c0 exp(g)."""
        w = self.w
        self.c = self.c0 * np.exp(np.asarray(g, dtype=np.float64))
        integ = self.integrator
        integ.set_node_stiffness(self.c)
        z = np.zeros((w.nnodes, 3))
        integ.create_buffers(w.nlist, w.n_neigh, w.bond_stiffness,
                             w.critical_stretch, None, z, z, z, z, z,
                             np.zeros(w.nnodes), None, 1, 1)
        integ.adjoint_start(self.steps, self.every)
        integ.run(self.steps, self.dmags, self.fmags)
        return integ.write(None, None, None, None, None, None, None,
                           None)[0]

    def backward(self, dL_du):
        """dL/dg for dL/du of the last :meth:`forward`."""
        if self.c is None:
            raise RuntimeError("forward() has not been called")
        return self.c * self.integrator.adjoint_backward(dL_du)

    def state(self):
        """(u, damage, nlist) after the last forward."""
        u, _, _, _, _, dmg, nl, _ = self.integrator.write(
            None, None, None, None, None, None, None, None)
        return u, dmg, nl


def _function():
    import torch

    class StiffnessField(torch.autograd.Function):
        """u_N(g) with the adjoint as its backward."""

        @staticmethod
        def forward(ctx, g, solver):
            u = solver.forward(g.detach().cpu().numpy())
            ctx.solver = solver
            return torch.from_numpy(u).to(device=g.device, dtype=g.dtype)

        @staticmethod
        def backward(ctx, grad_u):
            dg = ctx.solver.backward(grad_u.detach().cpu().numpy())
            return torch.from_numpy(dg).to(device=grad_u.device,
                                           dtype=grad_u.dtype), None

    return StiffnessField


_FN = None


def displacements(g, solver):
    """u after the solver's steps as a differentiable function of the
    torch tensor g (float64, (nnodes,))."""
    global _FN
    if _FN is None:
        _FN = _function()
    return _FN.apply(g, solver)


def sensor_loss(g, solver, sensors, target):
    """Sum of squared errors of u at the sensor nodes against target
    (len(sensors), 3)."""
    u = displacements(g, solver)
    return ((u[sensors] - target) ** 2).sum()
