"""Built-in set-up predicates, evaluated in bulk on the GPU.

PeriPy builds bond types, boundary conditions, densities and tips by
calling Python predicates once per bond or per node (model.py:657-780,
1093-1212).  The predicates here are ordinary callables with the
reference's signatures -- ``is_bond_type(x, y)`` and ``is_*(x)`` -- so they
can be passed to :class:`~.model.Model` (or to PeriPy) unchanged; ``Model``
recognises them and evaluates every bond or node at once in libpd_hip
(``pd_plane_bond_types``, ``pd_first_box``), on the GPU when one is present
and on the host otherwise, with results equal to the per-bond and per-node
calls.
"""
import numpy as np

from . import _lib
from ._lib import check, dptr, iptr

__all__ = ["PlaneBondTypes", "Boxes"]


def _device(device):
    if device is not None:
        return int(device)
    return 0 if _lib.device_available(0) else -1


class PlaneBondTypes:
    """``is_bond_type`` of a body split by the plane ``x[axis] = position``.

    A bond whose two ends lie below the plane (``x[axis] < position``) has
    type ``below``, one whose ends both lie at or above it type ``above``,
    and a bond across it type ``across`` -- a two-material composite with
    an interface type, as in the composite configuration (material A below
    y = 0, B above, interface bonds across).
    """

    def __init__(self, axis, position, below=0, above=1, across=2,
                 device=None):
        if axis not in (0, 1, 2):
            raise ValueError("axis must be 0, 1 or 2")
        self.axis = int(axis)
        self.position = float(position)
        self.below, self.above, self.across = int(below), int(above), \
            int(across)
        self.device = device

    def __call__(self, x, y):
        """The per-bond predicate (model.py:735-737)."""
        bx = x[self.axis] < self.position
        by = y[self.axis] < self.position
        if bx and by:
            return self.below
        if not bx and not by:
            return self.above
        return self.across

    def bond_types(self, coords, nlist, n_neigh):
        """(nnodes, max_neighbours) intc bond types of every live slot
        (``n_neigh[i]`` first slots of row i), 0 elsewhere: the array the
        reference's per-bond loop builds (model.py:737-766)."""
        coords = _lib.as_f64(coords, None, "coords")
        nlist = np.ascontiguousarray(nlist, dtype=np.int32)
        n_neigh = np.ascontiguousarray(n_neigh, dtype=np.int32)
        n, k = nlist.shape
        out = np.zeros((n, k), dtype=np.int32)
        check(_lib.lib().pd_plane_bond_types(
            _device(self.device), n, k, dptr(coords), iptr(nlist),
            iptr(n_neigh), self.axis, self.position, self.below, self.above,
            self.across, iptr(out)))
        return out.astype(np.intc)


class Boxes:
    """A per-node predicate from axis-aligned boxes.

    ``boxes`` is a sequence of ``(lo, hi, value)``: a node x lies in a box
    when ``lo[d] <= x[d] < hi[d]`` for d = 0, 1, 2 (bounds may be
    ``-inf``/``inf``), and the first box that holds it gives its value;
    nodes in no box get ``default``.  The value is what the per-node
    predicate returns: for ``is_displacement_boundary`` and
    ``is_force_boundary`` a list of three floats or None, for ``is_tip`` a
    list of three tip names, tuples of names or None, for ``is_density`` a
    float.  Called with one coordinate it is that per-node predicate; called
    with a (3, M) array of coordinates (component first) it evaluates all M
    at once, the box of every node found in libpd_hip.
    """

    vectorised = True

    def __init__(self, boxes, default=None, device=None):
        self.lo = np.array([b[0] for b in boxes], dtype=np.float64).reshape(
            -1, 3)
        self.hi = np.array([b[1] for b in boxes], dtype=np.float64).reshape(
            -1, 3)
        self.values = [b[2] for b in boxes]
        self.default = default
        self.device = device
        vals = self.values + ([default] if default is not None else [])
        self.scalar = bool(vals) and all(
            isinstance(v, (int, float, np.integer, np.floating))
            for v in vals)
        if not self.scalar:
            for v in vals:
                if len(v) != 3:
                    raise ValueError("box values must be lists of length 3 "
                                     "(one per DOF) or floats")
        # boundary values: numbers or None; tips: anything else
        self.numeric = self.scalar or all(
            e is None or isinstance(e, (int, float, np.integer, np.floating))
            for v in vals for e in v)

    def _value(self, b):
        return self.default if b < 0 else self.values[b]

    def box_of(self, coords):
        """Index of the first box holding each node ((M, 3) coordinates),
        -1 for none."""
        x = _lib.as_f64(coords, None, "coords")
        n = x.shape[0]
        out = np.empty(n, dtype=np.int32)
        check(_lib.lib().pd_first_box(
            _device(self.device), n, dptr(x), self.lo.shape[0],
            dptr(self.lo), dptr(self.hi), iptr(out)))
        return out

    def __call__(self, x):
        x = np.asarray(x, dtype=np.float64)
        if x.ndim == 1:
            inside = np.all((self.lo <= x) & (x < self.hi), axis=1)
            hit = np.nonzero(inside)[0]
            v = self._value(int(hit[0]) if hit.size else -1)
            if self.scalar:
                return float(v)
            return [None] * 3 if v is None else list(v)
        box = self.box_of(np.ascontiguousarray(x.T))
        m = box.shape[0]
        if self.scalar:
            table = np.array([float(v) for v in self.values]
                             + [float(self.default)
                                if self.default is not None else np.nan])
            return table[box]
        rows = self.values + [self.default if self.default is not None
                              else [None] * 3]
        if self.numeric:
            table = np.array([[np.nan if e is None else float(e) for e in v]
                              for v in rows]).reshape(-1, 3)
            return np.ascontiguousarray(table[box].T)
        table = np.empty((len(rows), 3), dtype=object)
        for r, v in enumerate(rows):
            for d in range(3):
                table[r, d] = v[d]
        out = np.empty((3, m), dtype=object)
        out[:, :] = table[box].T
        return out
