"""Launch-shape autotuner for the step kernel.

The step kernel has two shape parameters (``Integrator.set_launch``):

* ``tile_nodes`` -- nodes per workgroup: the tile whose halo (the nodes and
  all their neighbours) is staged in LDS; bigger tiles stage fewer halo
  entries per node but hold more LDS;
* ``lanes_per_node`` -- the neighbour tile: how many lanes share one node's
  bond list (each lane takes 4-bond chunks round-robin), trading parallelism
  against the cross-lane reduction and idle lanes on short rows.

The best pair depends on the mesh (bond counts, boundary fraction, number of
tiles against the 256 CUs), so it is measured: each candidate is timed over
a few steps from the same initial state, and the fastest is kept.

PeriPy's OpenCL kernels give each node a workgroup and tile its neighbour
list over the workgroup's work-items (cl/peridynamics.cl:134-262); a search
over "nodes per workgroup" of 1-4 and "neighbour tiles" of 32-128 bonds
belongs to that shape.  Here a workgroup holds a tile of 64-256 nodes whose
halo it stages once, and a node's neighbour list is split over
``lanes_per_node`` lanes in rounds of 4 * lanes_per_node bonds, so the
searched sets are tile_nodes (nodes per workgroup) in {64, 128, 256} and
lanes_per_node in {1, 2, 4, 8} (neighbour tiles of 4 to 32 bonds).
"""
import itertools

DEFAULT_TILES = (64, 128, 256)
DEFAULT_LANES = (1, 2, 4, 8)


def candidates(tiles=DEFAULT_TILES, lanes=DEFAULT_LANES, min_threads=128,
               max_threads=1024):
    """(lanes_per_node, tile_nodes) pairs with a workgroup size in range."""
    return [(lp, t) for t, lp in itertools.product(tiles, lanes)
            if min_threads <= lp * t <= max_threads]


def autotune(apply, time_steps, cands=None, repeats=1):
    """Pick the fastest launch shape.

    ``apply(lanes_per_node, tile_nodes)`` sets the shape and restores the
    initial state; ``time_steps()`` runs the timed steps and returns
    milliseconds.  Returns ``(best, table)`` with ``table`` a list of
    ``((lanes_per_node, tile_nodes), ms)``; the best shape is applied (with
    the initial state restored) before returning.
    """
    cands = list(cands or candidates())
    table = []
    for lp, t in cands:
        best_ms = None
        for _ in range(max(1, repeats)):
            apply(lp, t)
            ms = time_steps()
            best_ms = ms if best_ms is None else min(best_ms, ms)
        table.append(((lp, t), best_ms))
    best = min(table, key=lambda e: e[1])[0]
    apply(*best)
    return best, table


def autotune_batch(batch, steps=10, cands=None, graph=True):
    """:func:`autotune` for a :class:`~.batch.SampleBatch`: each candidate
    re-tiles the batch and is timed over ``steps`` steps from the initial
    state, captured in one hipGraph (``graph``) as the batch's runs are, or
    as plain launches."""
    def apply(lp, t):
        batch.set_launch(lp, t)

    def time_steps():
        ms = batch.run(steps, graph=graph)
        batch.integrator.synchronize()
        return ms

    return autotune(apply, time_steps, cands)
