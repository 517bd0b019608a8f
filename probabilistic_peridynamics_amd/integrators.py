"""Integrators: the HIP drop-in for peripy/integrators.py.

Each integrator owns one simulation handle of libpd_hip (one GPU) and keeps
the reference's interface (peripy/integrators.py:11-300):

* ``__init__(dt, context=None)`` (dynamic integrators: ``(damping, dt,
  context=None)``), attribute ``context``;
* ``build(nnodes, degrees_freedom, max_neighbours, coords, volume, family,
  bc_types, bc_values, force_bc_types, force_bc_values,
  stiffness_corrections, bond_types, densities)`` -- called by
  ``Model.__init__`` (model.py:475-479);
* ``create_buffers(nlist, n_neigh, bond_stiffness, critical_stretch, plus_cs,
  u, ud, udd, force, body_force, damage, regimes, nregimes, nbond_types)`` --
  called by ``Model._simulate_initialise`` (model.py:1569-1571);
* ``__call__(displacement_bc_magnitude, force_bc_magnitude)`` -- one step;
* ``write(u, ud, udd, force, body_force, damage, nlist, n_neigh)`` -- damage
  and device-to-host copies, returning the 8 arrays in that order;
* ``_build_special()``, ``_create_special_buffers()``.

Classes:

* :class:`Euler` -- the reference's Cython ``Euler`` (integrators.py:303-459):
  PMB only, ``context`` is None so :class:`~.model.Model` builds the Cython
  neighbour-list layout (width ``family.max()``, broken bonds removed by
  compaction, peridynamics.pyx:102-156), breaking on ``!(|s| < s0)``.
* :class:`EulerHIP`, :class:`EulerCromerHIP`, :class:`VelocityVerletHIP` --
  the OpenCL integrators (integrators.py:462-764): ``context`` is a
  :class:`HIPContext`, so Model builds the power-of-two, ``-1``-marked layout;
  they accept stiffness corrections, bond types and n-linear damage models.
  ``EulerCL``, ``EulerCromerCL`` and ``VelocityVerletCL`` are aliases.

All steps run on the GPU; there is no CPU fallback.
"""
from abc import ABC, abstractmethod
import ctypes

import numpy as np

from . import _lib
from ._lib import (PD_EULER, PD_EULER_CROMER, PD_VELOCITY_VERLET,
                   PD_LAYOUT_COMPACT, PD_LAYOUT_SLOT, check, dptr, iptr)


class HIPContext:
    """One HIP device: the ``context`` of the HIP integrators.

    Plays the role of the pyopencl context of the reference
    (peripy/cl/utilities.py:28-44): it names the device the integrator runs
    on.  It is always truthy.
    """

    def __init__(self, device=0):
        self.device = int(device)

    def __bool__(self):
        return True

    def __repr__(self):
        return f"HIPContext(device={self.device})"

    @property
    def devices(self):
        return [self.device]


def get_context(device=0):
    """A :class:`HIPContext` for ``device``, or None if there is no GPU.

    Replaces get_context (peripy/cl/utilities.py:28-44), which returns None
    when no suitable device exists.
    """
    try:
        n = _lib.device_count()
    except OSError:
        return None
    if device < 0 or device >= n:
        return None
    return HIPContext(device)


def output_device_info(context):
    """A one-line description of the context's device (cl/utilities.py:47)."""
    name, arch = _lib.device_name(context.device)
    return f"HIP device {context.device}: {name} ({arch})"


class ContextError(Exception):
    """No suitable context was found by :func:`get_context`."""

    def __init__(self):
        message = ("No suitable context was found. You can manually specify "
                   "the context by passing it to the integrator with the "
                   "'context' argument.")
        super().__init__(message)


class Integrator(ABC):
    """Base class for the HIP integrators (integrators.py:11-300)."""

    _kind = PD_EULER
    _layout = PD_LAYOUT_SLOT

    @abstractmethod
    def __init__(self, dt, context=None):
        self.dt = dt
        if context is None:
            self.context = get_context()
            if not isinstance(self.context, HIPContext):
                raise ContextError
        else:
            if not isinstance(context, HIPContext):
                raise TypeError("context must be a HIPContext object")
            self.context = context
        self._device = self.context.device
        self._sim = None
        self.damping = getattr(self, "damping", 0.0)

    # -- reference interface -------------------------------------------
    @abstractmethod
    def __call__(self, displacement_bc_magnitude, force_bc_magnitude):
        """Conduct one step of the integrator."""

    @abstractmethod
    def _build_special(self):
        """Checks special to the integrator (integrators.py:69-74)."""

    @abstractmethod
    def _create_special_buffers(self):
        """Buffers special to the integrator (integrators.py:76-82)."""

    def build(self, nnodes, degrees_freedom, max_neighbours, coords, volume,
              family, bc_types, bc_values, force_bc_types, force_bc_values,
              stiffness_corrections, bond_types, densities):
        """Upload the static arrays (integrators.py:84-188)."""
        self.nnodes = int(nnodes)
        self.degrees_freedom = int(degrees_freedom)
        self.max_neighbours = int(max_neighbours)
        self.densities = densities
        self._build_special()
        n, k = self.nnodes, self.max_neighbours
        self.coords = _lib.as_f64(coords, (n, 3), "coords")
        self.volume = _lib.as_f64(volume, (n,), "volume")
        self.family = _lib.as_i32(family, (n,), "family")
        self.bc_types = _lib.as_i32(bc_types, (n, 3), "bc_types")
        self.bc_values = _lib.as_f64(bc_values, (n, 3), "bc_values")
        self.force_bc_types = _lib.as_i32(force_bc_types, (n, 3),
                                          "force_bc_types")
        self.force_bc_values = _lib.as_f64(force_bc_values, (n, 3),
                                           "force_bc_values")
        self.stiffness_corrections = _lib.as_f64(
            stiffness_corrections, (n, k), "stiffness_corrections")
        self.bond_types = _lib.as_i32(bond_types, (n, k), "bond_types")
        dens = _lib.as_f64(densities, (n, 3), "densities")
        self._free()
        sim = ctypes.c_void_p()
        check(_lib.lib().pd_build(
            ctypes.byref(sim), self._device, self._kind, self._layout,
            float(self.dt), float(self.damping), n, k, dptr(self.coords),
            dptr(self.volume), iptr(self.family), iptr(self.bc_types),
            dptr(self.bc_values), iptr(self.force_bc_types),
            dptr(self.force_bc_values), dptr(self.stiffness_corrections),
            iptr(self.bond_types), dptr(dens)))
        self._sim = sim

    def create_buffers(self, nlist, n_neigh, bond_stiffness,
                       critical_stretch, plus_cs, u, ud, udd, force,
                       body_force, damage, regimes, nregimes, nbond_types):
        """Upload the simulation state (integrators.py:190-254)."""
        self._need_sim()
        n, k = self.nnodes, self.max_neighbours
        nlist = _lib.as_i32(nlist, (n, k), "nlist")
        n_neigh = _lib.as_i32(n_neigh, (n,), "n_neigh")
        nregimes = int(nregimes)
        nbond_types = int(nbond_types)
        if nregimes == 1 and nbond_types == 1:
            ks = np.array([np.asarray(bond_stiffness, dtype=np.float64)
                           .reshape(-1)[0]])
            ss = np.array([np.asarray(critical_stretch, dtype=np.float64)
                           .reshape(-1)[0]])
            pcs = None
            regimes = None
        else:
            ks = _lib.as_f64(np.reshape(bond_stiffness, -1))
            ss = _lib.as_f64(np.reshape(critical_stretch, -1))
            pcs = _lib.as_f64(np.reshape(plus_cs, -1))
            if ks.size != nregimes * nbond_types or ss.size != ks.size \
                    or pcs.size != ks.size:
                raise ValueError("damage model tables must hold "
                                 "nbond_types*nregimes values")
            regimes = (None if regimes is None else
                       _lib.as_i32(regimes, (n, k), "regimes"))
        u = _lib.as_f64(u, (n, 3), "u")
        ud = _lib.as_f64(ud, (n, 3), "ud")
        udd = _lib.as_f64(udd, (n, 3), "udd")
        check(_lib.lib().pd_create_buffers(
            self._sim, iptr(nlist), iptr(n_neigh), dptr(ks), dptr(ss),
            dptr(pcs), nregimes, nbond_types, dptr(u), dptr(ud), dptr(udd),
            iptr(regimes)))
        self.nregimes = nregimes
        self.nbond_types = nbond_types
        self._create_special_buffers()

    def write(self, u, ud, udd, force, body_force, damage, nlist, n_neigh):
        """Damage, then device-to-host copies (integrators.py:285-300).

        Each array argument is filled in place when it is a C-contiguous
        array of the right shape and dtype, otherwise a new array is
        returned in its place.
        """
        self._need_sim()
        n, k = self.nnodes, self.max_neighbours

        def out(a, shape, dtype):
            if (isinstance(a, np.ndarray) and a.shape == shape
                    and a.dtype == dtype and a.flags["C_CONTIGUOUS"]
                    and a.flags["WRITEABLE"]):
                return a
            return np.empty(shape, dtype=dtype)

        u = out(u, (n, 3), np.float64)
        ud = out(ud, (n, 3), np.float64)
        udd = out(udd, (n, 3), np.float64)
        force = out(force, (n, 3), np.float64)
        body_force = out(body_force, (n, 3), np.float64)
        damage = out(damage, (n,), np.float64)
        nlist = out(nlist, (n, k), np.int32)
        n_neigh = out(n_neigh, (n,), np.int32)
        check(_lib.lib().pd_write(
            self._sim, dptr(u), dptr(ud), dptr(udd), dptr(force),
            dptr(body_force), dptr(damage), iptr(nlist), iptr(n_neigh)))
        return (u, ud, udd, force, body_force, damage, nlist, n_neigh)

    # -- extensions ----------------------------------------------------
    def write_fields(self, u, damage):
        """:meth:`write` of u and damage only (no neighbour-list export or
        copy): what a periodic mesh file needs."""
        self._need_sim()
        n = self.nnodes
        u = u if (isinstance(u, np.ndarray) and u.shape == (n, 3)
                  and u.dtype == np.float64 and u.flags["C_CONTIGUOUS"]
                  and u.flags["WRITEABLE"]) else np.empty((n, 3))
        damage = damage if (isinstance(damage, np.ndarray)
                            and damage.shape == (n,)
                            and damage.dtype == np.float64
                            and damage.flags["C_CONTIGUOUS"]
                            and damage.flags["WRITEABLE"]) else np.empty(n)
        check(_lib.lib().pd_write(self._sim, dptr(u), None, None, None, None,
                                  dptr(damage), None, None))
        return u, damage

    def run(self, steps, displacement_bc_magnitudes=None,
            force_bc_magnitudes=None):
        """``steps`` calls of ``__call__`` in one library call."""
        self._need_sim()
        d = self._mags(displacement_bc_magnitudes, steps)
        f = self._mags(force_bc_magnitudes, steps)
        check(_lib.lib().pd_run(self._sim, int(steps), dptr(d), dptr(f)))

    def timed_run(self, steps, displacement_bc_magnitudes=None,
                  force_bc_magnitudes=None):
        """:meth:`run` timed with HIP events on the library's stream (ms)."""
        self._need_sim()
        d = self._mags(displacement_bc_magnitudes, steps)
        f = self._mags(force_bc_magnitudes, steps)
        ms = ctypes.c_float(0.0)
        check(_lib.lib().pd_run_timed(self._sim, int(steps), dptr(d),
                                      dptr(f), ctypes.byref(ms)))
        return float(ms.value)

    def run_graph(self, steps, displacement_bc_magnitudes=None,
                  force_bc_magnitudes=None):
        """:meth:`run` captured as one hipGraph and launched once; returns
        the graph's run time in ms (HIP events on the library's stream)."""
        self._need_sim()
        d = self._mags(displacement_bc_magnitudes, steps)
        f = self._mags(force_bc_magnitudes, steps)
        ms = ctypes.c_float(0.0)
        check(_lib.lib().pd_run_graph(self._sim, int(steps), dptr(d),
                                      dptr(f), ctypes.byref(ms)))
        return float(ms.value)

    def adjoint_start(self, steps, checkpoint_every=100):
        """Record the next ``steps`` steps (run by ``__call__`` / ``run``)
        for :meth:`adjoint_backward`: break steps and a checkpoint of u
        every ``checkpoint_every`` steps.  Euler with a node stiffness
        field only."""
        self._need_sim()
        check(_lib.lib().pd_adjoint_start(self._sim, int(steps),
                                          int(checkpoint_every)))

    def adjoint_backward(self, dL_du):
        """dL/dc of the node stiffness field for dL/du (nnodes, 3) after the
        recorded steps (the break mask of the recording held fixed)."""
        self._need_sim()
        g = _lib.as_f64(dL_du, (self.nnodes, 3), "dL_du")
        out = np.zeros(self.nnodes)
        check(_lib.lib().pd_adjoint_backward(self._sim, dptr(g), dptr(out)))
        return out

    def adjoint_backward_device(self, dL_du_ptr, dL_dc_ptr):
        """:meth:`adjoint_backward` between device buffers of this
        integrator's GPU, given by address: dL_du (nnodes, 3) f64 in,
        dL_dc (nnodes,) f64 out, both in the caller's node order and
        C-contiguous.  dL_du must be complete before the call; dL_dc is
        complete on return."""
        self._need_sim()
        check(_lib.lib().pd_adjoint_backward_device(
            self._sim, ctypes.c_void_p(int(dL_du_ptr)),
            ctypes.c_void_p(int(dL_dc_ptr))))

    def set_node_stiffness_device(self, c_ptr):
        """:meth:`set_node_stiffness` from a device buffer (nnodes,) f64 of
        this integrator's GPU, given by address; copied before return."""
        self._need_sim()
        check(_lib.lib().pd_set_node_stiffness_device(
            self._sim, ctypes.c_void_p(int(c_ptr))))

    def write_device(self, u_ptr=None, damage_ptr=None):
        """u (nnodes, 3) and damage (nnodes,) f64 into device buffers of
        this integrator's GPU, given by address (0/None to skip)."""
        self._need_sim()
        check(_lib.lib().pd_write_device(
            self._sim, ctypes.c_void_p(int(u_ptr or 0)) if u_ptr else None,
            ctypes.c_void_p(int(damage_ptr)) if damage_ptr else None))

    def adjoint_stop(self):
        self._need_sim()
        check(_lib.lib().pd_adjoint_stop(self._sim))

    def set_groups(self, group):
        """Per-node tiling groups (tiles never span two groups), or None.
        Takes effect at the next ``create_buffers``."""
        self._need_sim()
        g = (None if group is None
             else _lib.as_i32(group, (self.nnodes,), "group"))
        check(_lib.lib().pd_set_groups(self._sim, iptr(g)))

    def set_tips(self, tip_types):
        """Tip groups for :meth:`monitor`: ``{name: [(node, dof), ...]}``
        (the ``tip_types`` of model.py:1093-1212), order kept."""
        self._need_sim()
        self._tip_names = list(tip_types)
        self._tips = {name: list(tip_types[name]) for name in tip_types}
        node, dof, group = [], [], []
        for g, name in enumerate(self._tip_names):
            for i, j in tip_types[name]:
                node.append(i)
                dof.append(j)
                group.append(g)
        arr = [np.ascontiguousarray(x, dtype=np.int32)
               for x in (node, dof, group)]
        check(_lib.lib().pd_set_tips(self._sim, len(node), *map(iptr, arr),
                                     len(self._tip_names)))

    _SUMS = ("displacement", "velocity", "acceleration", "force",
             "body_force")

    def monitor(self):
        """The sums one periodic write records, computed on the device:
        ``(model, tips)`` with ``model`` = {displacement, velocity,
        acceleration, force, body_force: sums of u, ud, udd, force*V,
        body_force*V over all DOFs; damage_sum} and ``tips`` = {tip name:
        the same five sums over its (node, dof) entries}."""
        self._need_sim()
        names = getattr(self, "_tip_names", [])
        model = np.zeros(6)
        tips = np.zeros(5 * max(len(names), 1))
        check(_lib.lib().pd_monitor(self._sim, dptr(model), dptr(tips)))
        out = dict(zip(self._SUMS, model[:5]))
        out["damage_sum"] = model[5]
        return out, {name: dict(zip(self._SUMS, tips[5 * g:5 * g + 5]))
                     for g, name in enumerate(names)}

    def synchronize(self):
        self._need_sim()
        check(_lib.lib().pd_synchronize(self._sim))

    def live_bonds(self):
        """Number of live directed bonds (sum of n_neigh) right now."""
        self._need_sim()
        c = ctypes.c_longlong(0)
        check(_lib.lib().pd_live_bonds(self._sim, ctypes.byref(c)))
        return int(c.value)

    def regimes(self):
        """(nnodes, max_neighbours) bond regimes (n-linear models)."""
        self._need_sim()
        out = np.zeros((self.nnodes, self.max_neighbours), dtype=np.int32)
        check(_lib.lib().pd_get_regimes(self._sim, iptr(out)))
        return out

    def set_node_stiffness(self, c):
        """Per-node stiffness field (k_ij = (c_i + c_j)/2), None to clear."""
        self._need_sim()
        c = None if c is None else _lib.as_f64(c, (self.nnodes,), "c")
        check(_lib.lib().pd_set_node_stiffness(self._sim, dptr(c)))

    def set_break_mode(self, tension_only):
        """Break on !(s < s0) (True) or !(|s| < s0) (False, default)."""
        self._need_sim()
        check(_lib.lib().pd_set_break_mode(self._sim, int(bool(tension_only))))

    def set_launch(self, lanes_per_node=4, tile_nodes=128):
        """Lanes per node and nodes per tile of the step kernel.

        ``tile_nodes`` takes effect at the next ``create_buffers``."""
        self._need_sim()
        check(_lib.lib().pd_set_launch(self._sim, int(lanes_per_node),
                                       int(tile_nodes)))

    def tile_stats(self):
        """dict(ntiles, hmax, hstride, halo_total, slots) of the tiling."""
        self._need_sim()
        nt, hm, hs = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(0)
        ht, sl = ctypes.c_longlong(0), ctypes.c_longlong(0)
        check(_lib.lib().pd_tile_stats(self._sim, ctypes.byref(nt),
                                       ctypes.byref(hm), ctypes.byref(hs),
                                       ctypes.byref(ht), ctypes.byref(sl)))
        return dict(ntiles=nt.value, hmax=hm.value, hstride=hs.value,
                    halo_total=ht.value, slots=sl.value)

    # -- internals -----------------------------------------------------
    @staticmethod
    def _mags(m, steps):
        if m is None:
            return None
        m = np.ascontiguousarray(m, dtype=np.float64).reshape(-1)
        if m.size < steps:
            raise ValueError("fewer magnitudes than steps")
        return m

    def _need_sim(self):
        if self._sim is None:
            raise RuntimeError("build() has not been called")

    def _free(self):
        if getattr(self, "_sim", None) is not None:
            _lib.lib().pd_destroy(self._sim)
            self._sim = None

    def __del__(self):
        try:
            self._free()
        except Exception:
            pass


class Euler(Integrator):
    r"""Euler integrator with the Cython path's semantics, on the GPU.

    ``u(t + dt) = u(t) + dt f(t)`` (integrators.py:303-459).  PMB only;
    ``context`` is None so that :class:`~.model.Model` lays the neighbour
    list out as the Cython path does (model.py:580-582), and broken bonds are
    removed from it by the Cython compaction (peridynamics.pyx:102-156).
    Each node's bond forces are summed in the order of
    peridynamics.pyx:34-99, so force and u equal the Cython path's bitwise.
    """

    _kind = PD_EULER
    _layout = PD_LAYOUT_COMPACT

    def __init__(self, dt, device=0, exact_order=True):
        """Create an :class:`Euler` integrator for HIP device ``device``.

        ``exact_order=False`` sums each node's bonds in the fast step's own
        order instead of the Cython path's: the same broken bonds, lists and
        damage, force and u to rounding (see :meth:`set_exact_order`).
        """
        self.dt = dt
        self.context = None
        self._device = int(device)
        self._sim = None
        self.damping = 0.0
        self.exact_order = bool(exact_order)

    def __call__(self, displacement_bc_magnitude, force_bc_magnitude):
        """One step: break bonds, bond force, Euler update."""
        self._need_sim()
        check(_lib.lib().pd_step(self._sim, float(displacement_bc_magnitude),
                                 float(force_bc_magnitude)))

    def build(self, nnodes, degrees_freedom, max_neighbours, coords, volume,
              family, bc_types, bc_values, force_bc_types, force_bc_values,
              stiffness_corrections, bond_types, densities):
        """integrators.py:382-423: no bond types, corrections or densities."""
        if bond_types is not None:
            raise ValueError("bond_types are not supported by this "
                             "integrator (expected {}, got {}), please use "
                             "EulerHIP instead".format(type(None),
                                                       type(bond_types)))
        if stiffness_corrections is not None:
            raise ValueError("stiffness_corrections are not supported by "
                             "this integrator (expected {}, got {}), please "
                             "use EulerHIP instead".format(
                                 type(None), type(stiffness_corrections)))
        if densities is not None:
            raise ValueError("densities are not supported by this "
                             "integrator (expected {}, got {}). This "
                             "integrator neglects inertial effects. Do not "
                             "supply a density or is_density argument or, "
                             "alternatively, please use a dynamic "
                             "integrator, such as EulerCromerHIP.".format(
                                 type(None), type(densities)))
        if _lib.device_count() <= self._device:
            raise ContextError
        super().build(nnodes, degrees_freedom, max_neighbours, coords, volume,
                      family, bc_types, bc_values, force_bc_types,
                      force_bc_values, None, None, None)
        self.set_exact_order(self.exact_order)

    def set_exact_order(self, on):
        """Exact-order step on (True, the default) or off: the exact-order
        step sums each node's bonds in the Cython path's order and
        reproduces its force and u bitwise; the fast step sums them in its
        own order and matches to rounding."""
        self._need_sim()
        check(_lib.lib().pd_set_exact_order(self._sim, int(bool(on))))
        self.exact_order = bool(on)

    def create_buffers(self, nlist, n_neigh, bond_stiffness,
                       critical_stretch, plus_cs, u, ud, udd, force,
                       body_force, damage, regimes, nregimes, nbond_types):
        """integrators.py:351-380: PMB (one regime, one bond type) only."""
        if nregimes != 1:
            raise ValueError("n-linear damage model's are not supported by "
                             "this integrator. Please supply just one "
                             "bond_stiffness.")
        if nbond_types != 1:
            raise ValueError("n-material composite models are not supported "
                             "by this integrator. Please supply just one "
                             "material type and bond_stiffness.")
        super().create_buffers(nlist, n_neigh, bond_stiffness,
                               critical_stretch, plus_cs, u, ud, udd, force,
                               body_force, damage, regimes, nregimes,
                               nbond_types)
        # the Cython Euler keeps the body_force array it is given and never
        # fills it (integrators.py:371-380, 456-459): write returns it as is
        self._body_force = (
            np.zeros((self.nnodes, 3)) if body_force is None else
            np.array(body_force, dtype=np.float64).reshape(self.nnodes, 3))

    def write(self, u, ud, udd, force, body_force, damage, nlist, n_neigh):
        """integrators.py:456-459: the state, with body_force as given to
        create_buffers."""
        out = super().write(u, ud, udd, force, body_force, damage, nlist,
                            n_neigh)
        out[4][...] = self._body_force
        return out

    def monitor(self):
        """:meth:`Integrator.monitor` with the body-force sums of the
        unchanged body_force array (numpy's sum for the model, entry order
        for the tips, as the reference's write loop computes them)."""
        model, tips = super().monitor()
        v = self.volume
        model["body_force"] = np.sum(self._body_force * v[:, np.newaxis])
        for name, entries in getattr(self, "_tips", {}).items():
            acc = 0.0
            for i, j in entries:
                acc += self._body_force[i, j] * v[i]
            tips[name]["body_force"] = acc
        return model, tips

    def _build_special(self):
        """There are no special programs for this integrator."""

    def _create_special_buffers(self):
        """There are no special buffers for this integrator."""


class EulerHIP(Integrator):
    r"""Euler integrator, slot layout (integrators.py:462-541).

    ``u(t + dt) = u(t) + dt f(t)``.
    """

    _kind = PD_EULER
    _layout = PD_LAYOUT_SLOT

    def __init__(self, *args, **kwargs):
        super().__init__(*args, **kwargs)

    def __call__(self, displacement_bc_magnitude, force_bc_magnitude):
        """One step: bond force with in-place breaking, Euler update."""
        self._need_sim()
        check(_lib.lib().pd_step(self._sim, float(displacement_bc_magnitude),
                                 float(force_bc_magnitude)))

    def _build_special(self):
        """integrators.py:505-519: densities are rejected."""
        if self.densities is not None:
            raise ValueError("densities are not supported by this "
                             "integrator (expected {}, got {}). This "
                             "integrator neglects inertial effects. Do not "
                             "supply a density or is_density argument or, "
                             "alternatively, use a dynamic integrator, "
                             "such as EulerCromerHIP.".format(
                                 type(None), type(self.densities)))

    def _create_special_buffers(self):
        """There are no special buffers for this integrator."""


class EulerCromerHIP(Integrator):
    r"""Euler-Cromer integrator with damping (integrators.py:544-648).

    ``udd = (f - eta ud)/rho; ud += dt udd; u += dt ud``
    (cl/euler_cromer.cl:29-34).
    """

    _kind = PD_EULER_CROMER
    _layout = PD_LAYOUT_SLOT

    def __init__(self, damping, *args, **kwargs):
        """:arg float damping: dynamic relaxation damping [kg/(m^3 s)]."""
        self.damping = damping
        super().__init__(*args, **kwargs)
        self.damping = damping

    def __call__(self, displacement_bc_magnitude, force_bc_magnitude):
        """One step: bond force with in-place breaking, Euler-Cromer."""
        self._need_sim()
        check(_lib.lib().pd_step(self._sim, float(displacement_bc_magnitude),
                                 float(force_bc_magnitude)))

    def _build_special(self):
        """integrators.py:608-620: densities are required."""
        if self.densities is None:
            raise ValueError(
                "densities must be supplied when using EulerCromerHIP "
                "integrator (got {}). This integrator is dynamic "
                " and requires the density or is_density argument to be "
                "supplied to :class:Model, alternatively, use a static "
                " integrator, such as EulerHIP.".format(type(self.densities)))

    def _create_special_buffers(self):
        """There are no special buffers for this integrator."""


class VelocityVerletHIP(Integrator):
    r"""Velocity-Verlet integrator with damping (integrators.py:651-764).

    cl/velocity_verlet.cl:29-35: ``ud1 = ud + dt/2 udd;
    udd1 = (f - eta ud1)/rho; ud = ud1 + dt/2 udd1; udd = udd1;
    u += dt (ud + dt/2 udd1)``.
    """

    _kind = PD_VELOCITY_VERLET
    _layout = PD_LAYOUT_SLOT

    def __init__(self, damping, *args, **kwargs):
        """:arg float damping: dynamic relaxation damping [kg/(m^3 s)]."""
        self.damping = damping
        super().__init__(*args, **kwargs)
        self.damping = damping

    def __call__(self, displacement_bc_magnitude, force_bc_magnitude):
        """One step: bond force with in-place breaking, Velocity-Verlet."""
        self._need_sim()
        check(_lib.lib().pd_step(self._sim, float(displacement_bc_magnitude),
                                 float(force_bc_magnitude)))

    def _build_special(self):
        """integrators.py:722-734: densities are required."""
        if self.densities is None:
            raise ValueError(
                "densities must be supplied when using VelocityVerletHIP "
                "integrator (got {}). This integrator is dynamic "
                " and requires the density or is_density argument to be "
                "supplied to :class:Model, alternatively, use a static "
                " integrator, such as EulerHIP.".format(type(self.densities)))

    def _create_special_buffers(self):
        """There are no special buffers for this integrator."""


# The reference's OpenCL class names, for drop-in use.
EulerCL = EulerHIP
EulerCromerCL = EulerCromerHIP
VelocityVerletCL = VelocityVerletHIP
