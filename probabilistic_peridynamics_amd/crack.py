"""Initial cracks: bulk crack predicates on the GPU and create_crack.

PeriPy builds an initial crack from a pairwise predicate called once per
bond in Python (``initial_crack_helper``, model.py:1578-1612), which takes
minutes on a million-node mesh.  :class:`LineCrack` is the bulk form of the
predicate PeriPy's tests and examples use (peripy/test/conftest.py:43-63): a
straight crack, evaluated for every bond at once by ``pd_line_crack_pairs``
on the GPU (or the same code on the host).  An instance is a valid
``initial_crack`` argument of :class:`~.model.Model`, and its
:meth:`LineCrack.predicate` is the pairwise form for
``initial_crack_helper``; both give the same pairs in the same order.
"""
import ctypes

import numpy as np

from . import _lib


def create_crack(crack, nlist, n_neigh):
    """Remove the bonds of ``crack`` ((ncrack, 2) pairs) from the neighbour
    list in place (create_crack.pyx:1-43): for each pair (i, j), j leaves
    row i and then i leaves row j -- the first slot holding it takes the
    row's last live entry, the vacated last slot becomes -1 and n_neigh
    drops by one.  nlist (int32, (N, K)) and n_neigh (int32, (N,)) must be
    C-contiguous; they are modified in place."""
    crack = np.ascontiguousarray(np.asarray(crack, dtype=np.int32)
                                 .reshape(-1, 2))
    for name, a, nd in (("nlist", nlist, 2), ("n_neigh", n_neigh, 1)):
        if (not isinstance(a, np.ndarray) or a.dtype != np.int32
                or a.ndim != nd or not a.flags["C_CONTIGUOUS"]
                or not a.flags.writeable):
            raise ValueError(f"{name} must be a writable C-contiguous int32 "
                             f"array of {nd} dimensions")
    n, k = nlist.shape
    if n_neigh.shape[0] != n:
        raise ValueError("n_neigh has {} entries for {} rows of nlist".format(
            n_neigh.shape[0], n))
    _lib.check(_lib.lib().pd_create_crack(
        n, k, crack.shape[0], _lib.iptr(crack), _lib.iptr(nlist),
        _lib.iptr(n_neigh)))


class LineCrack:
    """A straight crack: the bonds that cross coordinate ``axis`` = ``split``
    and whose line meets coordinate ``axis`` = ``at`` at a ``cross_axis``
    coordinate strictly between ``lo`` and ``hi``.

    PeriPy's test crack (conftest.py:43-63, crack length 0.3 across the
    middle of the unit plate) is
    ``LineCrack(0, 1, split=0.5 + 1e-6, at=0.5, lo=0.5 * (1 - 0.3),
    hi=0.5 * (1 + 0.3))``.

    ``device``: the GPU that searches the bonds, -1 for the host, None for
    GPU 0 when present and the host otherwise.
    """

    def __init__(self, axis, cross_axis, split, at, lo, hi, device=None):
        if {axis, cross_axis} - {0, 1, 2} or axis == cross_axis:
            raise ValueError("axis and cross_axis must be two of 0, 1, 2")
        self.axis = int(axis)
        self.cross_axis = int(cross_axis)
        self.split = float(split)
        self.at = float(at)
        self.lo = float(lo)
        self.hi = float(hi)
        self.device = device

    def predicate(self, x, y):
        """The pairwise form: 1 when the bond (x, y) crosses the crack."""
        a, c = self.axis, self.cross_axis
        p1, p2 = (y, x) if x[a] > y[a] else (x, y)
        if p1[a] < self.split and p2[a] > self.split:
            m = (p2[c] - p1[c]) / (p2[a] - p1[a])
            b = p1[c] - m * p1[a]
            height = m * self.at + b
            if height > self.lo and height < self.hi:
                return 1
        return 0

    def __call__(self, coords, nlist, n_neigh):
        """(npairs, 2) int32 crack pairs (i, j), i < j, in the order of
        initial_crack_helper: i ascending, then i's row order."""
        coords = _lib.as_f64(coords)
        nlist = _lib.as_i32(nlist)
        n_neigh = _lib.as_i32(n_neigh)
        n, k = nlist.shape
        if coords.shape != (n, 3) or n_neigh.shape != (n,):
            raise ValueError("coords must be (N, 3) and n_neigh (N,) for an "
                             "(N, K) nlist")
        device = self.device
        if device is None:
            device = 0 if _lib.device_available(0) else -1
        args = (int(device), n, k, _lib.dptr(coords), _lib.iptr(nlist),
                _lib.iptr(n_neigh), self.axis, self.cross_axis, self.split,
                self.at, self.lo, self.hi)
        count = ctypes.c_longlong(0)
        _lib.check(_lib.lib().pd_line_crack_pairs(*args, None, 0,
                                                  ctypes.byref(count)))
        pairs = np.empty((count.value, 2), dtype=np.int32)
        if count.value:
            _lib.check(_lib.lib().pd_line_crack_pairs(
                *args, _lib.iptr(pairs), count.value, ctypes.byref(count)))
        return pairs
