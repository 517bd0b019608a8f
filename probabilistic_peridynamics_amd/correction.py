"""Stiffness-correction factors: the functions of peripy/correction.pyx.

Same names, arguments and in-place behaviour as the reference's Cython
module (correction.pyx:5-277), with one extra keyword, ``device``: the GPU
that computes the factors (``pd_set_*`` of libpd_hip, one thread per bond
slot), ``-1`` for the same code on the host, or None (default) for GPU 0
when one is present and the host otherwise.

Every function multiplies its first argument, an (nnodes, max_neighbours)
float64 array, in place: for each live bond of a pair i < j, slot (i, s)
and the first slot of row j that holds i are multiplied by the pair's
factor -- the loop of the reference, bit for bit.  As with the typed
memoryviews of the Cython functions, arrays of the wrong dtype raise
ValueError ("Buffer dtype mismatch").
"""
import numpy as np

from . import _lib

__all__ = ["set_imprecise_surface_correction",
           "set_precise_surface_correction", "set_volume_correction",
           "set_micromodulus_function"]

_CTYPE = {np.dtype(np.float64): "double", np.dtype(np.intc): "int"}


def _buffer(a, dtype, ndim, name):
    """A C-contiguous view (or copy) of ``a`` after the checks a Cython
    typed memoryview makes."""
    dtype = np.dtype(dtype)
    if not isinstance(a, np.ndarray):
        a = np.asarray(a)
    if a.dtype != dtype:
        raise ValueError("Buffer dtype mismatch, expected '{}' but got '{}' "
                         "({})".format(_CTYPE[dtype], a.dtype, name))
    if a.ndim != ndim:
        raise ValueError("Buffer has wrong number of dimensions (expected {},"
                         " got {}) ({})".format(ndim, a.ndim, name))
    return np.ascontiguousarray(a)


def _device(device):
    if device is None:
        return 0 if _lib.device_available(0) else -1
    return int(device)


def _inplace(values, call):
    """Run ``call`` on a C-contiguous float64 array holding ``values`` and
    leave the result in ``values``."""
    if not isinstance(values, np.ndarray) or values.dtype != np.float64:
        raise ValueError("Buffer dtype mismatch, expected 'double' but got "
                         "'{}'".format(getattr(values, "dtype", type(values))))
    if values.ndim != 2:
        raise ValueError("Buffer has wrong number of dimensions (expected 2,"
                         " got {})".format(values.ndim))
    if not values.flags.writeable:
        raise ValueError("buffer source array is read-only")
    work = values if values.flags["C_CONTIGUOUS"] else \
        np.ascontiguousarray(values)
    _lib.check(call(work))
    if work is not values:
        values[...] = work


def _graph(nlist, n_neigh, values):
    nlist = _buffer(nlist, np.intc, 2, "nlist")
    n_neigh = _buffer(n_neigh, np.intc, 1, "n_neigh")
    n, k = nlist.shape
    if n_neigh.shape[0] != n:
        raise ValueError("n_neigh has {} entries for {} rows of nlist".format(
            n_neigh.shape[0], n))
    if getattr(values, "shape", None) != (n, k):
        raise ValueError("the correction array must have the shape of nlist "
                         "{} (got {})".format((n, k),
                                              getattr(values, "shape", None)))
    return nlist, n_neigh, n, k


def set_imprecise_surface_correction(stiffness_corrections, nlist, n_neigh,
                                     average_volume, family_volume_bulk,
                                     device=None):
    """correction.pyx:5-56: the volume method with family volumes
    n_neigh[i] * average_volume."""
    nlist, n_neigh, n, k = _graph(nlist, n_neigh, stiffness_corrections)
    dev = _device(device)
    _inplace(stiffness_corrections,
             lambda w: _lib.lib().pd_set_imprecise_surface_correction(
                 dev, n, k, _lib.dptr(w), _lib.iptr(nlist),
                 _lib.iptr(n_neigh), float(average_volume),
                 float(family_volume_bulk)))


def set_precise_surface_correction(stiffness_corrections, nlist, n_neigh,
                                   volume, family_volume_bulk, device=None):
    """correction.pyx:59-120: the volume method with family volumes summed
    from the nodal volumes."""
    nlist, n_neigh, n, k = _graph(nlist, n_neigh, stiffness_corrections)
    volume = _buffer(volume, np.float64, 1, "volume")
    if volume.shape[0] != n:
        raise ValueError("volume has {} entries for {} nodes".format(
            volume.shape[0], n))
    dev = _device(device)
    _inplace(stiffness_corrections,
             lambda w: _lib.lib().pd_set_precise_surface_correction(
                 dev, n, k, _lib.dptr(w), _lib.iptr(nlist),
                 _lib.iptr(n_neigh), _lib.dptr(volume),
                 float(family_volume_bulk)))


def _coords(r0, n):
    r0 = _buffer(r0, np.float64, 2, "r0")
    if r0.shape != (n, 3):
        raise ValueError("r0 shape is wrong (expected {}, got {})".format(
            (n, 3), r0.shape))
    return r0


def set_volume_correction(volume_corrections, r0, nlist, n_neigh, horizon,
                          node_radius, volume_correction, device=None):
    """correction.pyx:123-205: partial-volume factors (cPA_HHB) when
    volume_correction == 0, else 1."""
    nlist, n_neigh, n, k = _graph(nlist, n_neigh, volume_corrections)
    r0 = _coords(r0, n)
    dev = _device(device)
    _inplace(volume_corrections,
             lambda w: _lib.lib().pd_set_volume_correction(
                 dev, n, k, _lib.dptr(w), _lib.dptr(r0), _lib.iptr(nlist),
                 _lib.iptr(n_neigh), float(horizon), float(node_radius),
                 int(volume_correction)))


def set_micromodulus_function(micromodulus_values, r0, nlist, n_neigh,
                              horizon, micromodulus_function, device=None):
    """correction.pyx:208-277: conical micromodulus values when
    micromodulus_function == 0, else 1."""
    nlist, n_neigh, n, k = _graph(nlist, n_neigh, micromodulus_values)
    r0 = _coords(r0, n)
    dev = _device(device)
    _inplace(micromodulus_values,
             lambda w: _lib.lib().pd_set_micromodulus_function(
                 dev, n, k, _lib.dptr(w), _lib.dptr(r0), _lib.iptr(nlist),
                 _lib.iptr(n_neigh), float(horizon),
                 int(micromodulus_function)))
