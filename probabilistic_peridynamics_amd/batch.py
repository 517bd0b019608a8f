"""Batched probabilistic samples: many independent meshes in one simulation.

PeriPy runs a probabilistic study as one ``Model.simulate`` call per sample
with new stiffness parameters (peripy/model.py:1523-1542,
examples/example2/example.py:361-369).  Here the samples' meshes are
concatenated into one node set (node ids offset per sample, so no bond joins
two samples), each sample is a tiling group (tiles never span samples), the
per-node stiffness field ``c`` gives each bond ``k_ij = (c_i + c_j)/2``, and
all samples advance together: one step kernel launch per step for the whole
batch, the steps of a run captured into one hipGraph and launched once.

Every node's arithmetic is the same as in a run of its sample alone with the
same lanes per node, so per-sample results are bitwise those of solo runs.
"""
import numpy as np

from . import integrators as it


def concatenate(samples):
    """The samples' arrays joined into one node set: node ids of sample s
    are offset by the sizes of samples 0..s-1, neighbour lists are padded
    with -1 to the widest, and ``group`` holds each node's sample index."""
    sizes = np.array([w.nnodes for w in samples])
    offsets = np.concatenate([[0], np.cumsum(sizes)])
    n = int(offsets[-1])
    k = max(w.max_neighbours for w in samples)
    cat = np.concatenate
    nlist = np.full((n, k), -1, dtype=np.int32)
    for w, o in zip(samples, offsets[:-1]):
        rows = nlist[o:o + w.nnodes, :w.max_neighbours]
        rows[:] = np.where(w.nlist >= 0, w.nlist + o, -1)
    return {
        "sizes": sizes, "offsets": offsets,
        "coords": cat([w.coords for w in samples]),
        "volume": cat([w.volume for w in samples]),
        "family": cat([w.family for w in samples]).astype(np.int32),
        "bc_types": cat([w.bc_types for w in samples]),
        "bc_values": cat([w.bc_values for w in samples]),
        "force_bc_types": cat([w.force_bc_types for w in samples]),
        "force_bc_values": cat([w.force_bc_values for w in samples]),
        "cnode": cat([w.extra["cnode"] for w in samples]),
        "group": np.repeat(np.arange(len(samples), dtype=np.int32), sizes),
        "nlist": nlist,
        "n_neigh": cat([w.n_neigh for w in samples]).astype(np.int32),
    }


class SampleBatch:
    """Run a list of Euler :class:`~.configs.Workload` samples together.

    Every sample needs ``extra["cnode"]`` (its per-node stiffness field) and
    the same ``dt``, ``critical_stretch`` and displacement schedule
    (``disp_rate``); bond breaking is the Cython criterion ``!(|s| < s0)``
    (``tension_only`` switches to the OpenCL one).
    """

    def __init__(self, samples, context=None, lanes_per_node=4,
                 tile_nodes=128, tension_only=False):
        if not samples:
            raise ValueError("no samples")
        w0 = samples[0]
        for w in samples:
            if w.integrator != "euler":
                raise ValueError("batched samples use the Euler integrator")
            if (w.dt != w0.dt or w.critical_stretch != w0.critical_stretch
                    or w.disp_rate != w0.disp_rate
                    or w.force_scale != w0.force_scale):
                raise ValueError("samples must share dt, critical_stretch "
                                 "and the boundary-condition schedule")
            if "cnode" not in w.extra:
                raise ValueError("every sample needs extra['cnode']")
        self.samples = samples
        cat = concatenate(samples)
        for name, v in cat.items():
            setattr(self, name, v)
        n, k = self.coords.shape[0], self.nlist.shape[1]
        self.nnodes, self.max_neighbours = n, k
        self.live0 = int(self.n_neigh.sum())
        self.dt = w0.dt
        self.critical_stretch = w0.critical_stretch
        self.bond_stiffness = w0.bond_stiffness
        self.disp_rate, self.force_scale = w0.disp_rate, w0.force_scale
        self.steps_done = 0

        self.integrator = it.EulerHIP(self.dt, context=context)
        self.integrator.build(n, 3, k, self.coords, self.volume, self.family,
                              self.bc_types, self.bc_values,
                              self.force_bc_types, self.force_bc_values,
                              None, None, None)
        self.integrator.set_groups(self.group)
        self.integrator.set_break_mode(tension_only)
        self.set_launch(lanes_per_node, tile_nodes)

    # -- state ---------------------------------------------------------------
    def set_launch(self, lanes_per_node, tile_nodes):
        """Launch shape; rebuilds the tiles and restarts from the initial
        state."""
        self.integrator.set_launch(lanes_per_node, tile_nodes)
        self.reset()

    def reset(self):
        """Back to the initial state (u = 0, every initial bond live)."""
        integ = self.integrator
        integ.set_node_stiffness(self.cnode)
        z = np.zeros((self.nnodes, 3))
        integ.create_buffers(self.nlist, self.n_neigh, self.bond_stiffness,
                             self.critical_stretch, None, z, z, z, z, z,
                             np.zeros(self.nnodes), None, 1, 1)
        self.steps_done = 0

    def magnitudes(self, steps):
        s = np.arange(self.steps_done + 1, self.steps_done + steps + 1,
                      dtype=np.float64)
        return self.disp_rate * s, np.full(steps, float(self.force_scale))

    # -- running -------------------------------------------------------------
    def run(self, steps, graph=True):
        """Advance every sample ``steps`` steps; returns the GPU time (ms)
        of the run (HIP events on the library's stream).  ``graph``: capture
        the steps into one hipGraph and launch it once, else one launch per
        step."""
        d, f = self.magnitudes(steps)
        if graph:
            ms = self.integrator.run_graph(steps, d, f)
        else:
            ms = self.integrator.timed_run(steps, d, f)
        self.steps_done += steps
        return ms

    def results(self):
        """Per-sample dicts of u, damage, nlist (sample-local ids, -1 for
        broken or padding) and n_neigh."""
        n, k = self.nnodes, self.max_neighbours
        u = np.zeros((n, 3))
        dmg = np.zeros(n)
        nl = np.zeros((n, k), dtype=np.int32)
        nn = np.zeros(n, dtype=np.int32)
        z = np.zeros((n, 3))
        self.integrator.write(u, z.copy(), z.copy(), z.copy(), z.copy(), dmg,
                              nl, nn)
        out = []
        for w, o in zip(self.samples, self.offsets[:-1]):
            sl = slice(o, o + w.nnodes)
            rows = nl[sl, :w.max_neighbours]
            out.append({"u": u[sl], "damage": dmg[sl],
                        "nlist": np.where(rows >= 0, rows - o, -1),
                        "n_neigh": nn[sl]})
        return out


def solo_integrator(w, context=None, lanes_per_node=4, tile_nodes=128,
                    tension_only=False):
    """One sample on its own simulation: the sequential (plain) baseline."""
    integ = it.EulerHIP(w.dt, context=context)
    integ.build(w.nnodes, 3, w.max_neighbours, w.coords, w.volume, w.family,
                w.bc_types, w.bc_values, w.force_bc_types, w.force_bc_values,
                None, None, None)
    integ.set_break_mode(tension_only)
    integ.set_launch(lanes_per_node, tile_nodes)
    integ.set_node_stiffness(w.extra["cnode"])
    z = np.zeros((w.nnodes, 3))
    integ.create_buffers(w.nlist, w.n_neigh, w.bond_stiffness,
                         w.critical_stretch, None, z, z, z, z, z,
                         np.zeros(w.nnodes), None, 1, 1)
    return integ
