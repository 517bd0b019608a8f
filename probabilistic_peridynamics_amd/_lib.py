"""ctypes binding of libpd_hip.so (declared in include/pd_hip.h).

The library is built in-tree by :func:`probabilistic_peridynamics_amd.build
.build_library` (``hipcc --offload-arch=gfx950``).  Loading it needs the HIP
runtime but not a GPU; every entry point that touches device memory fails with
a clear error when no GPU is present.  There is no fallback: if the library is
missing, :func:`lib` raises.
"""
import ctypes
import pathlib

import numpy as np

PACKAGE_DIR = pathlib.Path(__file__).resolve().parent
LIB_PATH = PACKAGE_DIR / "libpd_hip.so"

PD_EULER, PD_EULER_CROMER, PD_VELOCITY_VERLET = 0, 1, 2
PD_LAYOUT_COMPACT, PD_LAYOUT_SLOT = 0, 1

PD_OK = 0
PD_E_ARG = -1
PD_E_HIP = -2
PD_E_NODEV = -3
PD_E_STATE = -4
PD_E_LIMIT = -5

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)
_F = ctypes.POINTER(ctypes.c_float)
_LL = ctypes.POINTER(ctypes.c_longlong)
_SIM = ctypes.c_void_p

# name -> (restype, argtypes); every symbol include/pd_hip.h declares
SIGNATURES = {
    "pd_last_error": (ctypes.c_char_p, []),
    "pd_device_count": (ctypes.c_int, [_I]),
    "pd_device_name": (ctypes.c_int, [ctypes.c_int, ctypes.c_char_p,
                                      ctypes.c_int, ctypes.c_char_p,
                                      ctypes.c_int]),
    "pd_build": (ctypes.c_int, [
        ctypes.POINTER(_SIM), ctypes.c_int, ctypes.c_int, ctypes.c_int,
        ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int,
        _D, _D, _I, _I, _D, _I, _D, _D, _I, _D]),
    "pd_create_buffers": (ctypes.c_int, [
        _SIM, _I, _I, _D, _D, _D, ctypes.c_int, ctypes.c_int, _D, _D, _D,
        _I]),
    "pd_set_node_stiffness": (ctypes.c_int, [_SIM, _D]),
    "pd_set_node_stiffness_device": (ctypes.c_int, [_SIM, ctypes.c_void_p]),
    "pd_set_break_mode": (ctypes.c_int, [_SIM, ctypes.c_int]),
    "pd_set_exact_order": (ctypes.c_int, [_SIM, ctypes.c_int]),
    "pd_set_launch": (ctypes.c_int, [_SIM, ctypes.c_int, ctypes.c_int]),
    "pd_set_groups": (ctypes.c_int, [_SIM, _I]),
    "pd_step": (ctypes.c_int, [_SIM, ctypes.c_double, ctypes.c_double]),
    "pd_run_graph": (ctypes.c_int, [_SIM, ctypes.c_int, _D, _D, _F]),
    "pd_adjoint_start": (ctypes.c_int, [_SIM, ctypes.c_int, ctypes.c_int]),
    "pd_adjoint_backward": (ctypes.c_int, [_SIM, _D, _D]),
    "pd_adjoint_backward_device": (ctypes.c_int, [_SIM, ctypes.c_void_p,
                                                  ctypes.c_void_p]),
    "pd_adjoint_stop": (ctypes.c_int, [_SIM]),
    "pd_run": (ctypes.c_int, [_SIM, ctypes.c_int, _D, _D]),
    "pd_run_timed": (ctypes.c_int, [_SIM, ctypes.c_int, _D, _D, _F]),
    "pd_write": (ctypes.c_int, [_SIM, _D, _D, _D, _D, _D, _D, _I, _I]),
    "pd_write_device": (ctypes.c_int, [_SIM, ctypes.c_void_p,
                                       ctypes.c_void_p]),
    "pd_set_tips": (ctypes.c_int, [_SIM, ctypes.c_int, _I, _I, _I,
                                   ctypes.c_int]),
    "pd_monitor": (ctypes.c_int, [_SIM, _D, _D]),
    "pd_monitor_noalloc": (ctypes.c_int, [_SIM, _D, _D]),
    "pd_get_regimes": (ctypes.c_int, [_SIM, _I]),
    "pd_live_bonds": (ctypes.c_int, [_SIM, _LL]),
    "pd_synchronize": (ctypes.c_int, [_SIM]),
    "pd_tile_stats": (ctypes.c_int, [_SIM, _I, _I, _I, _LL, _LL]),
    "pd_destroy": (None, [_SIM]),
    "pd_neighbour_list_device": (ctypes.c_int, [
        ctypes.c_int, ctypes.c_int, _D, ctypes.c_double, _I, _I, ctypes.c_int,
        ctypes.c_int]),
    "pd_bond_check": (ctypes.c_int, [
        ctypes.c_int, ctypes.c_int, _D, _D, _D, _D, ctypes.c_double,
        ctypes.c_int, _D, _D, _D, _D, _I]),
    "pd_neighbour_list": (ctypes.c_int, [
        ctypes.c_int, _D, ctypes.c_double, _I, _I, ctypes.c_int,
        ctypes.c_int]),
    "pd_set_micromodulus_function": (ctypes.c_int, [
        ctypes.c_int, ctypes.c_int, ctypes.c_int, _D, _D, _I, _I,
        ctypes.c_double, ctypes.c_int]),
    "pd_set_volume_correction": (ctypes.c_int, [
        ctypes.c_int, ctypes.c_int, ctypes.c_int, _D, _D, _I, _I,
        ctypes.c_double, ctypes.c_double, ctypes.c_int]),
    "pd_set_imprecise_surface_correction": (ctypes.c_int, [
        ctypes.c_int, ctypes.c_int, ctypes.c_int, _D, _I, _I,
        ctypes.c_double, ctypes.c_double]),
    "pd_set_precise_surface_correction": (ctypes.c_int, [
        ctypes.c_int, ctypes.c_int, ctypes.c_int, _D, _I, _I, _D,
        ctypes.c_double]),
    "pd_line_crack_pairs": (ctypes.c_int, [
        ctypes.c_int, ctypes.c_int, ctypes.c_int, _D, _I, _I, ctypes.c_int,
        ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
        ctypes.c_double, _I, ctypes.c_longlong, _LL]),
    "pd_plane_bond_types": (ctypes.c_int, [
        ctypes.c_int, ctypes.c_int, ctypes.c_int, _D, _I, _I, ctypes.c_int,
        ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int, _I]),
    "pd_first_box": (ctypes.c_int, [
        ctypes.c_int, ctypes.c_int, _D, ctypes.c_int, _D, _D, _I]),
    "pd_create_crack": (ctypes.c_int, [
        ctypes.c_int, ctypes.c_int, ctypes.c_longlong, _I, _I, _I]),
}

_lib = None


class PDError(RuntimeError):
    """A libpd_hip call failed."""

    def __init__(self, code, message):
        super().__init__(f"libpd_hip error {code}: {message}")
        self.code = code


def lib():
    """Load libpd_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(
                f"{LIB_PATH} is missing: build it with "
                "`python -c 'import __graft_entry__ as g; g.build()'` or "
                "`python -m probabilistic_peridynamics_amd.build`")
        L = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(code):
    """Raise :class:`PDError` (or ValueError for argument errors)."""
    if code != PD_OK:
        msg = lib().pd_last_error().decode(errors="replace")
        if code == PD_E_ARG:
            raise ValueError(msg)
        raise PDError(code, msg)


def dptr(a):
    """Pointer to a C-contiguous float64 array; None stays NULL."""
    if a is None:
        return None
    if a.dtype != np.float64 or not a.flags["C_CONTIGUOUS"]:
        raise ValueError("Buffer dtype mismatch, expected 'double' "
                         "C-contiguous")
    return a.ctypes.data_as(_D)


def iptr(a):
    """Pointer to a C-contiguous int32 array; None stays NULL."""
    if a is None:
        return None
    if a.dtype != np.int32 or not a.flags["C_CONTIGUOUS"]:
        raise ValueError("Buffer dtype mismatch, expected 'int' "
                         "C-contiguous")
    return a.ctypes.data_as(_I)


def as_f64(a, shape=None, name="array"):
    """C-contiguous float64 copy-or-view of ``a``, shape-checked."""
    if a is None:
        return None
    out = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None and out.shape != tuple(shape):
        raise ValueError(f"{name} shape is wrong (expected {tuple(shape)}, "
                         f"got {out.shape})")
    return out


def as_i32(a, shape=None, name="array"):
    """C-contiguous int32 copy-or-view of ``a``, shape-checked."""
    if a is None:
        return None
    out = np.ascontiguousarray(a, dtype=np.int32)
    if shape is not None and out.shape != tuple(shape):
        raise ValueError(f"{name} shape is wrong (expected {tuple(shape)}, "
                         f"got {out.shape})")
    return out


def device_count():
    """Number of visible HIP devices (0 when none)."""
    c = ctypes.c_int(0)
    check(lib().pd_device_count(ctypes.byref(c)))
    return c.value


def device_name(device=0):
    """(marketing name, gfx architecture) of a device."""
    name = ctypes.create_string_buffer(256)
    arch = ctypes.create_string_buffer(256)
    check(lib().pd_device_name(device, name, 256, arch, 256))
    return name.value.decode(), arch.value.decode()


def device_available(device=0):
    """Whether HIP device ``device`` exists (False without the runtime)."""
    try:
        return 0 <= int(device) < device_count()
    except (OSError, PDError):
        return False


def _nl_call(n, coords, horizon, family, nlist, k, pad, device):
    if device is None:
        return lib().pd_neighbour_list(n, dptr(coords), float(horizon),
                                       iptr(family), iptr(nlist), k, pad)
    return lib().pd_neighbour_list_device(int(device), n, dptr(coords),
                                          float(horizon), iptr(family),
                                          iptr(nlist), k, pad)


def family_count(coords, horizon, device=None):
    """Number of nodes within ``horizon`` of each node (family); on the host
    (``device=None``) or on that GPU."""
    coords = as_f64(coords)
    family = np.zeros(coords.shape[0], dtype=np.int32)
    check(_nl_call(coords.shape[0], coords, horizon, family, None, 0, 0,
                   device))
    return family


def neighbour_list(coords, horizon, max_neighbours=None, pad_value=0,
                   device=None):
    """(family, nlist) of a point cloud; rows ascending, padded.  On the host
    (``device=None``) or on that GPU, with the same result.

    ``max_neighbours`` defaults to ``family.max()``.
    """
    coords = as_f64(coords)
    n = coords.shape[0]
    family = np.zeros(n, dtype=np.int32)
    if max_neighbours is None:
        family = family_count(coords, horizon, device)
        max_neighbours = int(family.max())
    nlist = np.empty((n, int(max_neighbours)), dtype=np.int32)
    check(_nl_call(n, coords, horizon, family, nlist, int(max_neighbours),
                   int(pad_value), device))
    return family, nlist


def bond_check(ri, rj, r0i, r0j, s0, tension_only=False, opencl=False,
               device=0):
    """The step kernel's stretch pipeline on a batch of bonds (GPU).

    ``opencl``: ri, rj are displacements and the stretch is the slot
    layout's OpenCL one.  Returns dict(l, l0, s, s_exact, broken)."""
    ri, rj, r0i, r0j = (as_f64(a) for a in (ri, rj, r0i, r0j))
    n = ri.shape[0]
    out = {k: np.empty(n) for k in ("l", "l0", "s", "s_exact")}
    out["broken"] = np.empty(n, dtype=np.int32)
    check(lib().pd_bond_check(
        device, n, dptr(ri), dptr(rj), dptr(r0i), dptr(r0j), float(s0),
        int(bool(tension_only)) | (2 if opencl else 0), dptr(out["l"]),
        dptr(out["l0"]),
        dptr(out["s"]), dptr(out["s_exact"]), iptr(out["broken"])))
    return out
