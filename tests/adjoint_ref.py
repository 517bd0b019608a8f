"""Reference gradient for the stiffness-field adjoint (test infrastructure).

A torch (CPU, float64) restatement of the Euler step with the Cython bond
arithmetic (peripy/peridynamics.pyx:34-99, 102-156, 159-185, with
r = r0 + u and spatial.pyx's stretch), a per-node stiffness field
c = c0 exp(g) with bond stiffness (c_i + c_j)/2, and the break mask held
fixed for differentiation: a bond breaks when !(|s| < s0) and stays broken,
the decision taken without gradient.  torch autograd through it gives the
reference gradient the GPU adjoint is checked against.
"""
import numpy as np
import torch


def edges(nlist):
    """Directed bonds (i, j) of a -1-padded neighbour list."""
    ii, kk = np.nonzero(nlist >= 0)
    return ii, nlist[ii, kk]


def forward(w, g, c0, steps, tension_only=False):
    """u after ``steps`` Euler steps from rest, and the final live mask (per
    directed bond of :func:`edges`)."""
    t = lambda a: torch.as_tensor(np.asarray(a, dtype=np.float64))
    ii, jj = edges(w.nlist)
    ii, jj = torch.as_tensor(ii), torch.as_tensor(jj)
    x0 = t(w.coords)
    vol = t(w.volume)
    free = t(w.bc_types == 0)
    bcv = t(w.bc_values)
    fbc = t(w.force_bc_values) * t(w.force_bc_types != 0)
    dmag, fmag = w.magnitudes(steps)
    c = c0 * torch.exp(g)
    k = (c[ii] + c[jj]) * 0.5
    e = x0[jj] - x0[ii]
    l0 = torch.sqrt((e[:, 0] * e[:, 0] + e[:, 1] * e[:, 1]) + e[:, 2] * e[:, 2])
    alive = torch.ones(len(ii), dtype=torch.bool)
    u = torch.zeros_like(x0)
    s0 = float(w.critical_stretch)
    for n in range(steps):
        r = x0 + u
        d = r[jj] - r[ii]
        l = torch.sqrt((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1])
                       + d[:, 2] * d[:, 2])
        s = (l - l0) / l0
        with torch.no_grad():
            ok = (s < s0) if tension_only else (torch.abs(s) < s0)
            alive = alive & ok
        fn = torch.where(alive, (s * k) / l, torch.zeros_like(s))
        f = (fn[:, None] * d) * vol[jj][:, None]
        F = torch.zeros_like(x0).index_add(0, ii, f) + fmag[n] * fbc
        u = torch.where(free > 0, u + w.dt * F, bcv * dmag[n])
    return u, alive


def loss_and_grad(w, g, c0, steps, sensors, target):
    """(L, dL/dg, u, alive) with L = sum of squared sensor errors."""
    gt = torch.as_tensor(np.asarray(g, dtype=np.float64)).requires_grad_()
    u, alive = forward(w, gt, c0, steps)
    L = ((u[sensors] - torch.as_tensor(target)) ** 2).sum()
    L.backward()
    return (float(L.detach()), gt.grad.numpy().copy(), u.detach().numpy(),
            alive.numpy())
