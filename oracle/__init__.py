"""Oracle: CPU restatement of PeriPy's time-stepping path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` import this package, and only as the
checker.  The product package ``probabilistic_peridynamics_amd`` never imports
it.

The arithmetic lives in ``pd_oracle.c`` (built by ``oracle/Makefile`` with the
reference's compile flags).  This module wraps it with ctypes and composes the
per-step drivers:

* :class:`CythonEuler` -- ``Euler.__call__`` of peripy/integrators.py:330-349:
  ``break_bonds(coords+u)`` then ``bond_force(coords+u)`` then
  ``update_displacement`` (peridynamics.pyx).  Bit-identical to the Cython
  path, including the compacted neighbour list and its stale tail slots.
* :class:`SlotStepper` -- the per-slot ("-1" breaking) path of
  ``EulerCL/EulerCromerCL/VelocityVerletCL.__call__``
  (integrators.py:484-503, 586-606, 700-720) with the kernels of
  peripy/cl/*.cl and their arithmetic as the OpenCL compiler that wrote the
  reference's fixtures evaluated it (contracted squared lengths and updates);
  ``arith=0`` gives the Cython stretch instead, for comparison.
"""
import ctypes
import pathlib
import subprocess

import numpy as np

_HERE = pathlib.Path(__file__).resolve().parent
_SO = _HERE / "liboracle.so"
_lib = None

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)


def build():
    """Compile ``liboracle.so`` with ``make`` (gcc, no GPU needed)."""
    subprocess.run(["make", "-s", "-C", str(_HERE)], check=True)


def lib():
    """Return the loaded oracle library, building it if it is missing."""
    global _lib
    if _lib is None:
        if not _SO.exists():
            build()
        L = ctypes.CDLL(str(_SO))
        sig = {
            "oracle_euclid": (ctypes.c_double, [_D, _D]),
            "oracle_strain": (ctypes.c_double, [_D, _D, _D, _D]),
            "oracle_strain2": (ctypes.c_double, [ctypes.c_double, _D, _D]),
            "oracle_damage": (None, [ctypes.c_int, _I, _I, _D]),
            "oracle_bond_force": (None, [
                ctypes.c_int, ctypes.c_int, _D, _D, _I, _I, _D,
                ctypes.c_double, _D, _I, ctypes.c_double, _D]),
            "oracle_break_bonds": (None, [
                ctypes.c_int, ctypes.c_int, _D, _D, _I, _I,
                ctypes.c_double]),
            "oracle_update_displacement": (None, [
                ctypes.c_int, _D, _D, _I, _D, ctypes.c_double,
                ctypes.c_double]),
            "oracle_create_crack": (None, [
                ctypes.c_int, _I, ctypes.c_int, _I, _I]),
            "oracle_slot_bond_force": (None, [
                ctypes.c_int, ctypes.c_int, _D, _D, _I, _D, ctypes.c_double,
                ctypes.c_double, ctypes.c_int, _D, _D, _I, _I, _D, _D, _D,
                ctypes.c_int, _I, _D, ctypes.c_double, _D, _D,
                ctypes.c_int]),
            "oracle_cl_strain": (ctypes.c_double, [_D, _D, _D, _D, _D, _D]),
            "oracle_update_euler_cl": (None, [
                ctypes.c_int, _D, _D, _I, _D, ctypes.c_double,
                ctypes.c_double]),
            "oracle_slot_damage": (None, [
                ctypes.c_int, ctypes.c_int, _I, _I, _I, _D]),
            "oracle_update_euler_cromer": (None, [
                ctypes.c_int, _D, _D, _D, _D, _I, _D, _D, ctypes.c_double,
                ctypes.c_double, ctypes.c_double]),
            "oracle_update_velocity_verlet": (None, [
                ctypes.c_int, _D, _D, _D, _D, _I, _D, _D, ctypes.c_double,
                ctypes.c_double, ctypes.c_double]),
            "oracle_positions": (None, [ctypes.c_int, _D, _D, _D]),
            "oracle_set_micromodulus_function": (None, [
                ctypes.c_int, ctypes.c_int, _D, _D, _I, _I, ctypes.c_double,
                ctypes.c_int]),
            "oracle_set_volume_correction": (None, [
                ctypes.c_int, ctypes.c_int, _D, _D, _I, _I, ctypes.c_double,
                ctypes.c_double, ctypes.c_int]),
            "oracle_set_imprecise_surface_correction": (None, [
                ctypes.c_int, ctypes.c_int, _D, _I, _I, ctypes.c_double,
                ctypes.c_double]),
            "oracle_set_precise_surface_correction": (None, [
                ctypes.c_int, ctypes.c_int, _D, _I, _I, _D, ctypes.c_double,
                _D]),
            "oracle_set_threads": (None, [ctypes.c_int]),
            "oracle_max_threads": (ctypes.c_int, []),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _d(a):
    """Pointer to a C-contiguous float64 array (None -> NULL)."""
    if a is None:
        return None
    if a.dtype != np.float64 or not a.flags["C_CONTIGUOUS"]:
        raise ValueError("Buffer dtype mismatch, expected 'double' "
                         "C-contiguous")
    return a.ctypes.data_as(_D)


def _i(a):
    """Pointer to a C-contiguous int32 array (None -> NULL)."""
    if a is None:
        return None
    if a.dtype != np.int32 or not a.flags["C_CONTIGUOUS"]:
        raise ValueError("Buffer dtype mismatch, expected 'int' "
                         "C-contiguous")
    return a.ctypes.data_as(_I)


# --------------------------------------------------------------------------
# Function-level mirrors of peripy.spatial / peripy.peridynamics
# --------------------------------------------------------------------------

def euclid(r1, r2):
    """spatial.pyx:6-31."""
    r1 = np.ascontiguousarray(r1, dtype=np.float64)
    r2 = np.ascontiguousarray(r2, dtype=np.float64)
    return lib().oracle_euclid(_d(r1), _d(r2))


def strain(r1, r2, r10, r20):
    """spatial.pyx:34-66."""
    a = [np.ascontiguousarray(x, dtype=np.float64) for x in (r1, r2, r10, r20)]
    return lib().oracle_strain(*[_d(x) for x in a])


def strain2(l, r10, r20):
    """spatial.pyx:69-97."""
    a = [np.ascontiguousarray(x, dtype=np.float64) for x in (r10, r20)]
    return lib().oracle_strain2(float(l), *[_d(x) for x in a])


def cl_strain(ui, uj, r0i, r0j):
    """cl/peridynamics.cl:65-75 as contracted: (s, l, l0) of one bond from
    the displacements and initial positions of its ends."""
    a = [np.ascontiguousarray(x, dtype=np.float64) for x in (ui, uj, r0i, r0j)]
    l, l0 = ctypes.c_double(0.0), ctypes.c_double(0.0)
    s = lib().oracle_cl_strain(*[_d(x) for x in a], ctypes.byref(l),
                               ctypes.byref(l0))
    return s, l.value, l0.value


def damage(n_neigh, family):
    """peridynamics.pyx:5-31."""
    out = np.empty(family.shape[0], dtype=np.float64)
    lib().oracle_damage(family.shape[0], _i(n_neigh), _i(family), _d(out))
    return out


def bond_force(r, r0, nlist, n_neigh, volume, bond_stiffness,
               force_bc_values, force_bc_types, force_bc_scale):
    """peridynamics.pyx:34-99 (returns a new force array)."""
    n, k = nlist.shape
    force = np.zeros((n, 3), dtype=np.float64)
    lib().oracle_bond_force(n, k, _d(r), _d(r0), _i(nlist), _i(n_neigh),
                            _d(volume), float(bond_stiffness),
                            _d(force_bc_values), _i(force_bc_types),
                            float(force_bc_scale), _d(force))
    return force


def break_bonds(r, r0, nlist, n_neigh, critical_strain):
    """peridynamics.pyx:102-156 (mutates nlist, n_neigh)."""
    n, k = nlist.shape
    lib().oracle_break_bonds(n, k, _d(r), _d(r0), _i(nlist), _i(n_neigh),
                             float(critical_strain))


def update_displacement(u, bc_values, bc_types, force, bc_scale, dt):
    """peridynamics.pyx:159-185 (mutates u)."""
    lib().oracle_update_displacement(u.shape[0], _d(u), _d(bc_values),
                                     _i(bc_types), _d(force),
                                     float(bc_scale), float(dt))


def create_crack(crack, nlist, n_neigh):
    """create_crack.pyx:1-43 (mutates nlist, n_neigh)."""
    crack = np.ascontiguousarray(np.asarray(crack, dtype=np.int32)
                                 .reshape(-1, 2))
    lib().oracle_create_crack(crack.shape[0], _i(crack), nlist.shape[1],
                              _i(nlist), _i(n_neigh))


# --------------------------------------------------------------------------
# Function-level mirrors of peripy.correction (all mutate their first array)
# --------------------------------------------------------------------------

def set_micromodulus_function(micromodulus_values, r0, nlist, n_neigh,
                              horizon, micromodulus_function):
    """correction.pyx:208-277."""
    n, k = nlist.shape
    r0 = np.ascontiguousarray(r0, dtype=np.float64)
    lib().oracle_set_micromodulus_function(
        n, k, _d(micromodulus_values), _d(r0), _i(nlist), _i(n_neigh),
        float(horizon), int(micromodulus_function))


def set_volume_correction(volume_corrections, r0, nlist, n_neigh, horizon,
                          node_radius, volume_correction):
    """correction.pyx:123-205."""
    n, k = nlist.shape
    r0 = np.ascontiguousarray(r0, dtype=np.float64)
    lib().oracle_set_volume_correction(
        n, k, _d(volume_corrections), _d(r0), _i(nlist), _i(n_neigh),
        float(horizon), float(node_radius), int(volume_correction))


def set_imprecise_surface_correction(stiffness_corrections, nlist, n_neigh,
                                     average_volume, family_volume_bulk):
    """correction.pyx:5-56."""
    n, k = nlist.shape
    lib().oracle_set_imprecise_surface_correction(
        n, k, _d(stiffness_corrections), _i(nlist), _i(n_neigh),
        float(average_volume), float(family_volume_bulk))


def set_precise_surface_correction(stiffness_corrections, nlist, n_neigh,
                                   volume, family_volume_bulk):
    """correction.pyx:59-120."""
    n, k = nlist.shape
    volume = np.ascontiguousarray(volume, dtype=np.float64)
    scratch = np.empty(n, dtype=np.float64)
    lib().oracle_set_precise_surface_correction(
        n, k, _d(stiffness_corrections), _i(nlist), _i(n_neigh), _d(volume),
        float(family_volume_bulk), _d(scratch))


# --------------------------------------------------------------------------
# Step drivers
# --------------------------------------------------------------------------

class CythonEuler:
    """The Cython ``Euler`` integrator (integrators.py:303-459).

    One call is ``break_bonds(coords+u)``, ``bond_force(coords+u)``,
    ``update_displacement`` (integrators.py:339-349, 433-453).  ``u``,
    ``nlist`` and ``n_neigh`` are the caller's arrays and are mutated.
    """

    def __init__(self, dt, coords, volume, family, bc_types, bc_values,
                 force_bc_types, force_bc_values, nlist, n_neigh,
                 bond_stiffness, critical_stretch, u=None):
        self.dt = float(dt)
        self.coords = np.ascontiguousarray(coords, dtype=np.float64)
        self.volume = np.ascontiguousarray(volume, dtype=np.float64)
        self.family = np.ascontiguousarray(family, dtype=np.int32)
        self.bc_types = np.ascontiguousarray(bc_types, dtype=np.int32)
        self.bc_values = np.ascontiguousarray(bc_values, dtype=np.float64)
        self.force_bc_types = np.ascontiguousarray(force_bc_types,
                                                   dtype=np.int32)
        self.force_bc_values = np.ascontiguousarray(force_bc_values,
                                                    dtype=np.float64)
        self.nlist = np.array(nlist, dtype=np.int32)
        self.n_neigh = np.array(n_neigh, dtype=np.int32)
        self.k = float(bond_stiffness)
        self.s0 = float(critical_stretch)
        n = self.coords.shape[0]
        self.u = (np.zeros((n, 3)) if u is None
                  else np.array(u, dtype=np.float64))
        self.force = np.zeros((n, 3))

    def __call__(self, displacement_bc_magnitude, force_bc_magnitude):
        break_bonds(self.coords + self.u, self.coords, self.nlist,
                    self.n_neigh, self.s0)
        self.force = bond_force(self.coords + self.u, self.coords,
                                self.nlist, self.n_neigh, self.volume,
                                self.k, self.force_bc_values,
                                self.force_bc_types, force_bc_magnitude)
        update_displacement(self.u, self.bc_values, self.bc_types,
                            self.force, displacement_bc_magnitude, self.dt)

    def damage(self):
        return damage(self.n_neigh, self.family)


EULER, EULER_CROMER, VELOCITY_VERLET = 0, 1, 2


class SlotStepper:
    """Per-slot ("-1") step: bond_force1..4 semantics + an integrator update.

    ``integrator``: 0 Euler (cl/euler.cl), 1 Euler-Cromer
    (cl/euler_cromer.cl), 2 Velocity-Verlet (cl/velocity_verlet.cl).
    ``arith``: 1 (default) the OpenCL kernels' arithmetic as compiled for the
    reference's fixtures -- xi = r0_j - r0_i, xi_eta = (u_j - u_i) + xi,
    squared lengths and updates contracted into fused multiply-adds; 0 the
    Cython stretch (r = r0 + u, spatial.pyx) and uncontracted updates.  Damage model: scalar ``bond_stiffness`` /
    ``critical_stretch`` (PMB), optionally a per-node stiffness field
    ``cnode`` (k_ij = (c_i + c_j)/2), optional ``stiffness_corrections``;
    or (nbond_types, nregimes) tables with ``bond_types`` and ``regimes``.
    All arrays are copied; state is in attributes.
    """

    def __init__(self, integrator, dt, coords, volume, family, bc_types,
                 bc_values, force_bc_types, force_bc_values, nlist,
                 bond_stiffness, critical_stretch, damping=0.0,
                 densities=None, stiffness_corrections=None, bond_types=None,
                 regimes=None, plus_cs=None, nregimes=1, cnode=None,
                 tension_only=False, u=None, ud=None, udd=None, arith=1):
        f64 = lambda a: None if a is None else np.array(a, dtype=np.float64)
        i32 = lambda a: None if a is None else np.array(a, dtype=np.int32)
        self.integrator = int(integrator)
        self.dt = float(dt)
        self.damping = float(damping)
        self.coords = f64(coords)
        n = self.coords.shape[0]
        self.volume = f64(volume)
        self.family = i32(family)
        self.bc_types = i32(bc_types)
        self.bc_values = f64(bc_values)
        self.force_bc_types = i32(force_bc_types)
        self.force_bc_values = f64(force_bc_values)
        self.nlist = i32(nlist)
        self.densities = f64(densities)
        self.corr = f64(stiffness_corrections)
        self.cnode = f64(cnode)
        self.tension_only = int(bool(tension_only))
        self.arith = int(arith)
        self.nregimes = int(nregimes)
        if np.ndim(bond_stiffness) == 0:
            self.k = float(bond_stiffness)
            self.s0 = float(critical_stretch)
            self.k_tab = self.s_tab = self.pc_tab = None
            self.bond_types = None
            self.regimes = None
        else:
            self.k = self.s0 = 0.0
            self.k_tab = f64(np.ravel(bond_stiffness))
            self.s_tab = f64(np.ravel(critical_stretch))
            self.pc_tab = f64(np.ravel(plus_cs))
            self.bond_types = (np.zeros_like(self.nlist) if bond_types is None
                               else i32(bond_types))
            self.regimes = (np.zeros_like(self.nlist) if regimes is None
                            else i32(regimes))
        z = lambda a: np.zeros((n, 3)) if a is None else f64(a)
        self.u, self.ud, self.udd = z(u), z(ud), z(udd)
        self.force = np.zeros((n, 3))
        self.body_force = np.zeros((n, 3))

    def __call__(self, displacement_bc_magnitude, force_bc_magnitude):
        L = lib()
        n, k = self.nlist.shape
        L.oracle_slot_bond_force(
            n, k, _d(self.u), _d(self.coords), _i(self.nlist),
            _d(self.volume), self.k, self.s0, self.tension_only,
            _d(self.cnode), _d(self.corr), _i(self.bond_types),
            _i(self.regimes), _d(self.k_tab), _d(self.s_tab),
            _d(self.pc_tab), self.nregimes, _i(self.force_bc_types),
            _d(self.force_bc_values), float(force_bc_magnitude),
            _d(self.force), _d(self.body_force), self.arith)
        if self.integrator == EULER:
            fn = (L.oracle_update_euler_cl if self.arith
                  else L.oracle_update_displacement)
            fn(n, _d(self.u), _d(self.bc_values), _i(self.bc_types),
               _d(self.force), float(displacement_bc_magnitude), self.dt)
        else:
            if not self.arith:
                raise ValueError("the dynamic integrators are restated with "
                                 "the OpenCL arithmetic only (arith=1)")
            fn = (L.oracle_update_euler_cromer
                  if self.integrator == EULER_CROMER
                  else L.oracle_update_velocity_verlet)
            fn(n, _d(self.force), _d(self.u), _d(self.ud), _d(self.udd),
               _i(self.bc_types), _d(self.bc_values), _d(self.densities),
               float(displacement_bc_magnitude), self.damping, self.dt)

    def connectivity(self):
        """(n_neigh, damage) recounted from the -1 slots."""
        n, k = self.nlist.shape
        n_neigh = np.empty(n, dtype=np.int32)
        dmg = np.empty(n, dtype=np.float64)
        lib().oracle_slot_damage(n, k, _i(self.nlist), _i(self.family),
                                 _i(n_neigh), _d(dmg))
        return n_neigh, dmg
