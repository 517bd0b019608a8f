"""Gradients of displacement losses with respect to the stiffness field.

The forward model is the Euler step of a :class:`~.configs.Workload` with a
per-node stiffness field ``c = c0 exp(g)`` (bond stiffness ``(c_i + c_j)/2``)
run on the GPU.  The backward pass is the library's hand-written adjoint
(``pd_adjoint_*``): the forward pass records the step at which each bond
broke and checkpoints ``u``; the backward pass replays each checkpoint
segment with the step kernel and runs the adjoint kernel back through it,
holding the recorded break mask fixed.  :class:`StiffnessField` wraps this
as a ``torch.autograd.Function`` of the log-stiffness ``g``; a ``g`` on the
GPU stays there (the field, u and the gradient pass to and from the library
by device address), a ``g`` on the host goes through numpy.
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
        # incremented by every forward: the recording a backward uses is
        # that of the last forward
        self.generation = 0

    def _restart(self):
        w = self.w
        z = np.zeros((w.nnodes, 3))
        self.integrator.create_buffers(w.nlist, w.n_neigh, w.bond_stiffness,
                                       w.critical_stretch, None, z, z, z, z,
                                       z, np.zeros(w.nnodes), None, 1, 1)

    def _record(self):
        integ = self.integrator
        integ.adjoint_start(self.steps, self.every)
        integ.run(self.steps, self.dmags, self.fmags)
        self.generation += 1

    def forward(self, g):
        """u after ``steps`` steps from rest, for c = c0 exp(g) (numpy)."""
        self.c = self.c0 * np.exp(np.asarray(g, dtype=np.float64))
        self.integrator.set_node_stiffness(self.c)
        self._restart()
        self._record()
        return self.integrator.write(None, None, None, None, None, None,
                                     None, None)[0]

    def backward(self, dL_du):
        """dL/dg for dL/du of the last :meth:`forward` (numpy)."""
        if self.c is None:
            raise RuntimeError("forward() has not been called")
        return self.c * self.integrator.adjoint_backward(dL_du)

    def forward_device(self, g):
        """:meth:`forward` for a float64 torch tensor g on this solver's
        GPU: c = c0 exp(g) is formed there and handed to the library by
        address, and u comes back as a tensor on the same GPU."""
        import torch
        g = g.detach().contiguous()
        c = self.c0 * torch.exp(g)
        torch.cuda.current_stream(g.device).synchronize()
        self._restart()
        self.integrator.set_node_stiffness_device(c.data_ptr())
        self.c = c
        self._record()
        u = torch.empty((self.w.nnodes, 3), dtype=torch.float64,
                        device=g.device)
        self.integrator.write_device(u.data_ptr(), None)
        return u

    def backward_device(self, dL_du):
        """:meth:`backward` for a float64 torch tensor dL/du (nnodes, 3) on
        this solver's GPU; dL/dg is returned on the same GPU."""
        import torch
        if self.c is None:
            raise RuntimeError("forward() has not been called")
        dL_du = dL_du.detach().contiguous()
        torch.cuda.current_stream(dL_du.device).synchronize()
        dc = torch.empty(self.w.nnodes, dtype=torch.float64,
                         device=dL_du.device)
        self.integrator.adjoint_backward_device(dL_du.data_ptr(),
                                                dc.data_ptr())
        c = self.c if torch.is_tensor(self.c) else torch.from_numpy(
            self.c).to(dL_du.device)
        return c * dc

    def state(self):
        """(u, damage, nlist) after the last forward."""
        u, _, _, _, _, dmg, nl, _ = self.integrator.write(
            None, None, None, None, None, None, None, None)
        return u, dmg, nl


def _function():
    import torch

    class StiffnessField(torch.autograd.Function):
        """u_N(g) with the adjoint as its backward.

        The solver holds one recording, that of its last forward.  A node
        of the graph whose recording was replaced by a later forward (two
        evaluations before one backward) re-runs its own forward first, so
        its gradient is always that of its own g."""

        @staticmethod
        def forward(ctx, g, solver):
            if g.dtype != torch.float64:
                raise TypeError("g must be float64")
            if g.is_cuda:
                u = solver.forward_device(g)
            else:
                u = torch.from_numpy(solver.forward(g.detach().numpy()))
            ctx.solver = solver
            ctx.generation = solver.generation
            ctx.save_for_backward(g)
            return u

        @staticmethod
        def backward(ctx, grad_u):
            solver = ctx.solver
            (g,) = ctx.saved_tensors
            if solver.generation != ctx.generation:
                # the recording belongs to another g: record this one again
                if g.is_cuda:
                    solver.forward_device(g)
                else:
                    solver.forward(g.detach().numpy())
            if grad_u.is_cuda:
                dg = solver.backward_device(grad_u.to(torch.float64))
            else:
                dg = torch.from_numpy(solver.backward(
                    grad_u.detach().to(torch.float64).numpy()))
            return dg.to(dtype=g.dtype), None

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
