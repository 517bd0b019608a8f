"""Synthetic workloads: jittered-lattice bars with their loading.

Each generator returns a :class:`Workload` -- the arrays an Integrator's
``build`` / ``create_buffers`` take (in the reference's formats) plus the
scalars of the run.  Nodes are cell centres of a regular lattice, x slowest
and z fastest, optionally jittered by a seeded uniform +-jitter*dx per
coordinate.  Volumes are dx^3.  The neighbour list is libpd_hip's cell-list
search (distance <= horizon), rows ascending.

Workloads (sizes as named):
  bar_euler:   0.4 x 0.08 x 0.08 m, dx = 0.01 (40 x 8 x 8 = 2560 nodes),
               horizon 3.01 dx, c = 18K/(pi horizon^4) with K = 1.5e10,
               s0 = 0.005, pre-crack through x = 0.2 for z < 0.04, x < 0.02
               clamped, x > 0.38 pulled +x by 1e-8 m per step, overdamped
               Euler dt = 1e-5.
  beam_vv:     2.0 x 0.2 x 0.2 m bar, dx = 0.2/47 (470 x 47 x 47 =
               1,038,230 nodes), jitter 0.25 dx (seed 1), horizon 3.01 dx,
               K padded to a power of two, rho = 2400, dt = 1e-8,
               Velocity-Verlet, damping 0, body force 1e4 N/m^3 in -z on
               the central strip |x - 1| < horizon, bottom nodes within a
               horizon of either end clamped in z.
  beam_euler:  the beam_vv mesh and loading in the compact layout, no
               densities, the Cython path's Euler with dt = 2e-16 (below
               its explicit stability limit, about 8e-16).
  composite_ec: the beam_vv geometry centred on y = 0, bond types A
               (y < 0), B (y >= 0) and interface; trilinear law with
               breakpoints (0.002, 0.004, 0.008) and slopes (c, 0.6c, 0.1c),
               c_B = 2.5 c_A, interface 0.5 c_A with A's breakpoints;
               Euler-Cromer, damping 2.5e6, rho 2400, dt 1e-8, ends pulled
               apart in x by 1e-9 m per step.
  uq_samples:  256 jittered 5:1:1 bars (jitter 0.25 dx, dx = 0.01) of
               N ~ logU[5000, 60000] nodes (sizes from default_rng(3),
               sample s from default_rng(1000 + s)); per-node stiffness
               c_i = c exp(g_i), g a Gaussian random field on the lattice
               (squared-exponential covariance, correlation length 5 dx,
               sigma 0.3); s0 = 0.005; the bar_euler loading (x < 2 dx
               clamped, x > L - 2 dx pulled +x by 1e-8 m per step);
               Euler dt = 1e-5.
"""
from dataclasses import dataclass, field

import numpy as np

from . import _lib

K_BULK = 1.5e10


def micromodulus(horizon, k_bulk=K_BULK):
    """c = 18 K / (pi horizon^4), the PMB bond stiffness."""
    return 18.0 * k_bulk / (np.pi * horizon ** 4)


def lattice(shape, dx, jitter=0.0, seed=None, origin=(0.0, 0.0, 0.0)):
    """Cell-centred lattice, x slowest, z fastest (float64, (N, 3))."""
    nx, ny, nz = shape
    g = np.stack(np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz),
                             indexing="ij"), -1).reshape(-1, 3)
    x = (g + 0.5) * dx + np.asarray(origin, dtype=np.float64)
    if jitter:
        rng = np.random.default_rng(seed)
        x = x + rng.uniform(-jitter * dx, jitter * dx, size=x.shape)
    return np.ascontiguousarray(x, dtype=np.float64)


@dataclass
class Workload:
    name: str
    coords: np.ndarray
    volume: np.ndarray
    family: np.ndarray
    nlist: np.ndarray
    n_neigh: np.ndarray
    horizon: float
    bond_stiffness: object
    critical_stretch: object
    dt: float
    integrator: str = "euler"          # euler | euler_cromer | velocity_verlet
    damping: float = 0.0
    densities: np.ndarray = None
    bc_types: np.ndarray = None
    bc_values: np.ndarray = None
    force_bc_types: np.ndarray = None
    force_bc_values: np.ndarray = None
    bond_types: np.ndarray = None
    plus_cs: np.ndarray = None
    stiffness_corrections: np.ndarray = None  # (N, K) factors, slot layout
    nregimes: int = 1
    nbond_types: int = 1
    disp_rate: float = 0.0             # displacement BC magnitude per step
    force_scale: float = 0.0           # force BC magnitude (every step)
    extra: dict = field(default_factory=dict)

    @property
    def nnodes(self):
        return self.coords.shape[0]

    @property
    def max_neighbours(self):
        return self.nlist.shape[1]

    def live_bonds(self):
        return int(self.n_neigh.sum())

    def magnitudes(self, steps, first_step=1):
        """(displacement, force) BC magnitudes of steps first..first+steps-1."""
        s = np.arange(first_step, first_step + steps, dtype=np.float64)
        return (self.disp_rate * s,
                np.full(steps, self.force_scale, dtype=np.float64))


def _search_device():
    """GPU 0 for the neighbour search when there is one, else the host."""
    return 0 if _lib.device_available(0) else None


def neighbours(coords, horizon, layout="slot"):
    """(family, nlist, n_neigh) in the reference's layouts: "slot" pads a
    power-of-two width with -1 (model.py:575-579), "compact" pads width
    family.max() with 0 (model.py:580-582).  The search runs on the GPU when
    there is one (the same result as the host search)."""
    device = _search_device()
    family = _lib.family_count(coords, horizon, device)
    if layout == "slot":
        k = 1 << (int(family.max()) - 1).bit_length()
        pad = -1
    else:
        k = int(family.max())
        pad = 0
    family, nlist = _lib.neighbour_list(coords, horizon, k, pad, device)
    return family, nlist, family.copy()


def remove_bonds(nlist, n_neigh, predicate, coords):
    """Crack: remove every bond (i, j), i < j, with predicate(x_i, x_j)
    true, by create_crack's swap-with-last / -1 rule (create_crack.pyx)."""
    n = coords.shape[0]
    pairs = []
    for i in range(n):
        row = nlist[i, :n_neigh[i]]
        js = row[row > i]
        if js.size:
            mask = predicate(coords[i][None, :], coords[js])
            pairs.extend((i, int(j)) for j in js[mask])
    from .model import create_crack
    create_crack(np.array(pairs, dtype=np.int64), nlist, n_neigh)
    return len(pairs)


def bar_euler(layout="slot", dt=1e-5, pull=1e-8):
    """The 2560-node Euler bar with a pre-crack."""
    dx = 0.01
    horizon = 3.01 * dx
    coords = lattice((40, 8, 8), dx)
    family, nlist, n_neigh = neighbours(coords, horizon, layout)

    def crack(xi, xj):
        lo = np.minimum(xi[:, 0], xj[:, 0])
        hi = np.maximum(xi[:, 0], xj[:, 0])
        return (lo < 0.2) & (hi > 0.2) & (xi[:, 2] < 0.04) & (xj[:, 2] < 0.04)

    ncrack = remove_bonds(nlist, n_neigh, crack, coords)
    n = coords.shape[0]
    bc_types = np.zeros((n, 3), dtype=np.int32)
    bc_values = np.zeros((n, 3))
    clamp = coords[:, 0] < 0.02
    bc_types[clamp] = 1
    pulled = coords[:, 0] > 0.38
    bc_types[pulled, 0] = 1
    bc_values[pulled, 0] = 1.0
    return Workload(
        name="bar_euler_2560", coords=coords, volume=np.full(n, dx ** 3),
        family=family, nlist=nlist, n_neigh=n_neigh, horizon=horizon,
        bond_stiffness=micromodulus(horizon), critical_stretch=0.005, dt=dt,
        integrator="euler", bc_types=bc_types, bc_values=bc_values,
        force_bc_types=np.zeros((n, 3), dtype=np.int32),
        force_bc_values=np.zeros((n, 3)), disp_rate=pull,
        extra={"crack_bonds": ncrack, "dx": dx})


def _bar_1m(shape, seed, origin=(0.0, 0.0, 0.0), layout="slot"):
    nx = shape[0]
    dx = 2.0 / nx
    horizon = 3.01 * dx
    coords = lattice(shape, dx, jitter=0.25, seed=seed, origin=origin)
    family, nlist, n_neigh = neighbours(coords, horizon, layout)
    return dx, horizon, coords, family, nlist, n_neigh


def beam_vv(shape=(470, 47, 47), seed=1, layout="slot"):
    """The 1,038,230-node Velocity-Verlet beam."""
    dx, horizon, coords, family, nlist, n_neigh = _bar_1m(shape, seed,
                                                          layout=layout)
    n = coords.shape[0]
    length = shape[0] * dx
    bc_types = np.zeros((n, 3), dtype=np.int32)
    support = (coords[:, 2] < horizon) & (
        (coords[:, 0] < horizon) | (coords[:, 0] > length - horizon))
    bc_types[support, 2] = 1
    fbc_types = np.zeros((n, 3), dtype=np.int32)
    fbc_values = np.zeros((n, 3))
    strip = np.abs(coords[:, 0] - length / 2) < horizon
    fbc_types[strip, 2] = 1
    fbc_values[strip, 2] = -1e4
    return Workload(
        name=f"beam_vv_{n}", coords=coords, volume=np.full(n, dx ** 3),
        family=family, nlist=nlist, n_neigh=n_neigh, horizon=horizon,
        bond_stiffness=micromodulus(horizon), critical_stretch=0.005,
        dt=1e-8, integrator="velocity_verlet", damping=0.0,
        densities=np.full((n, 3), 2400.0), bc_types=bc_types,
        bc_values=np.zeros((n, 3)), force_bc_types=fbc_types,
        force_bc_values=fbc_values, force_scale=1.0,
        extra={"dx": dx, "shape": shape, "seed": seed})


def beam_euler(shape=(470, 47, 47), seed=1):
    """The 1,038,230-node beam of beam_vv stepped by the Cython path's
    Euler integrator (compact layout): same mesh, supports and body-force
    strip, no densities.  Explicit Euler (u += dt F) is stable for
    dt < 2 / lambda_max of the bond stiffness operator; twice the largest
    eigenvalue of a node's block sum_j c V_j e e^T / l over the beam's nodes
    bounds lambda_max at 2.4e15, so dt < 8.3e-16, and the step is 2e-16 s
    (5e-15 s, past that limit, shatters the beam within tens of steps)."""
    w = beam_vv(shape, seed, layout="compact")
    w.name = f"beam_euler_{w.nnodes}"
    w.integrator = "euler"
    w.densities = None
    w.dt = 2e-16
    w.extra["layout"] = "compact"
    return w


def trilinear_tables(c_a):
    """(nbond_types=3, nregimes=3) tables of the composite trilinear law and
    their plus_cs (model.py:1050-1070's recurrence)."""
    c = np.array([c_a, 2.5 * c_a, 0.5 * c_a])
    k = np.stack([c, 0.6 * c, 0.1 * c], axis=1)
    s = np.tile(np.array([0.002, 0.004, 0.008]), (3, 1))
    pcs = np.zeros((3, 3))
    for r in range(1, 3):
        pcs[:, r] = pcs[:, r - 1] + k[:, r - 1] * s[:, r - 1] \
            - k[:, r] * s[:, r - 1]
    return k, s, pcs


def composite_ec(shape=(470, 47, 47), seed=2):
    """The 1M-node composite trilinear Euler-Cromer bar."""
    dx0 = 2.0 / shape[0]
    dx, horizon, coords, family, nlist, n_neigh = _bar_1m(
        shape, seed, origin=(0.0, -shape[1] * dx0 / 2, 0.0))
    n = coords.shape[0]
    k, s, pcs = trilinear_tables(micromodulus(horizon))
    side = (coords[:, 1] >= 0).astype(np.int32)
    jj = np.where(nlist >= 0, nlist, 0)
    si, sj = side[:, None], side[jj]
    bond_types = np.where(si == sj, si, 2).astype(np.int32)
    bond_types[nlist < 0] = 0
    length = shape[0] * dx
    bc_types = np.zeros((n, 3), dtype=np.int32)
    bc_values = np.zeros((n, 3))
    left = coords[:, 0] < horizon
    right = coords[:, 0] > length - horizon
    bc_types[left | right, 0] = 1
    bc_values[left, 0] = -1.0
    bc_values[right, 0] = 1.0
    return Workload(
        name=f"composite_ec_{n}", coords=coords, volume=np.full(n, dx ** 3),
        family=family, nlist=nlist, n_neigh=n_neigh, horizon=horizon,
        bond_stiffness=k, critical_stretch=s, plus_cs=pcs, nregimes=3,
        nbond_types=3, bond_types=bond_types, dt=1e-8,
        integrator="euler_cromer", damping=2.5e6,
        densities=np.full((n, 3), 2400.0), bc_types=bc_types,
        bc_values=bc_values, force_bc_types=np.zeros((n, 3), np.int32),
        force_bc_values=np.zeros((n, 3)), disp_rate=1e-9,
        extra={"dx": dx, "shape": shape, "seed": seed})


def make_integrator(w, context=None, lanes_per_node=None, tile_nodes=None):
    """A built integrator with the workload's state uploaded (launch shape
    optional: see Integrator.set_launch)."""
    from . import integrators as it
    if w.integrator == "euler" and w.extra.get("layout") == "compact":
        integ = it.Euler(w.dt)
    elif w.integrator == "euler":
        integ = it.EulerHIP(w.dt, context=context)
    elif w.integrator == "euler_cromer":
        integ = it.EulerCromerHIP(w.damping, w.dt, context=context)
    else:
        integ = it.VelocityVerletHIP(w.damping, w.dt, context=context)
    integ.build(w.nnodes, 3, w.max_neighbours, w.coords, w.volume, w.family,
                w.bc_types, w.bc_values, w.force_bc_types, w.force_bc_values,
                w.stiffness_corrections, w.bond_types, w.densities)
    if lanes_per_node is not None or tile_nodes is not None:
        integ.set_launch(lanes_per_node or 4, tile_nodes or 128)
    n = w.nnodes
    z = np.zeros((n, 3))
    integ.create_buffers(w.nlist, w.n_neigh, w.bond_stiffness,
                         w.critical_stretch, w.plus_cs, z, z, z, z, z,
                         np.zeros(n), None, w.nregimes, w.nbond_types)
    return integ


def gaussian_random_field(shape, dx, corr_len, sigma, rng):
    """Stationary Gaussian field on a lattice of ``shape`` cells with
    covariance sigma^2 exp(-|d|^2 / (2 corr_len^2)), by circulant embedding:
    the covariance is laid on a periodic grid padded by 4 correlation
    lengths per axis (where it has decayed to exp(-8)), its FFT gives the
    eigenvalues (negative round-off clipped), and white noise is coloured by
    their square roots."""
    pad = int(np.ceil(4.0 * corr_len / dx))
    m = [n + pad for n in shape]
    axes = []
    for md in m:
        k = np.arange(md)
        axes.append(np.minimum(k, md - k) * dx)
    d2 = (axes[0][:, None, None] ** 2 + axes[1][None, :, None] ** 2
          + axes[2][None, None, :] ** 2)
    cov = sigma ** 2 * np.exp(-d2 / (2.0 * corr_len ** 2))
    lam = np.clip(np.fft.fftn(cov).real, 0.0, None)
    w = rng.standard_normal(m)
    z = np.fft.ifftn(np.sqrt(lam) * np.fft.fftn(w)).real
    return z[:shape[0], :shape[1], :shape[2]]


def uq_sizes(nsamples=256, seed=3, nmin=5000, nmax=60000):
    """Bar shapes (5m, m, m) with 5 m^3 ~ N, N log-uniform in [nmin, nmax]."""
    rng = np.random.default_rng(seed)
    n = np.exp(rng.uniform(np.log(nmin), np.log(nmax), nsamples))
    m = np.maximum(2, np.rint(np.cbrt(n / 5.0))).astype(int)
    return [(5 * int(k), int(k), int(k)) for k in m]


def uq_sample(shape, s, dx=0.01, corr=5.0, sigma=0.3):
    """Sample s of the UQ batch on a (5m, m, m) bar (see module doc)."""
    rng = np.random.default_rng(1000 + s)
    horizon = 3.01 * dx
    coords = lattice(shape, dx, jitter=0.25, seed=rng)
    g = gaussian_random_field(shape, dx, corr * dx, sigma, rng).reshape(-1)
    c = micromodulus(horizon)
    family, nlist, n_neigh = neighbours(coords, horizon, "slot")
    n = coords.shape[0]
    length = shape[0] * dx
    bc_types = np.zeros((n, 3), dtype=np.int32)
    bc_values = np.zeros((n, 3))
    clamp = coords[:, 0] < 2 * dx
    bc_types[clamp] = 1
    pulled = coords[:, 0] > length - 2 * dx
    bc_types[pulled, 0] = 1
    bc_values[pulled, 0] = 1.0
    return Workload(
        name=f"uq_sample_{s}_{n}", coords=coords, volume=np.full(n, dx ** 3),
        family=family, nlist=nlist, n_neigh=n_neigh, horizon=horizon,
        bond_stiffness=c, critical_stretch=0.005, dt=1e-5,
        integrator="euler", bc_types=bc_types, bc_values=bc_values,
        force_bc_types=np.zeros((n, 3), dtype=np.int32),
        force_bc_values=np.zeros((n, 3)), disp_rate=1e-8,
        extra={"dx": dx, "shape": shape, "cnode": c * np.exp(g), "g": g})


def uq_samples(nsamples=256, seed=3, nmin=5000, nmax=60000):
    """The batch of UQ samples (list of Workloads, see module doc)."""
    return [uq_sample(shape, s)
            for s, shape in enumerate(uq_sizes(nsamples, seed, nmin, nmax))]
