"""Model: the reference's peridynamics model API over the HIP integrators.

Same constructor and ``simulate`` signatures as peripy/model.py:42-51 and
:1214-1218.  Set-up (mesh reading, volumes, neighbour list, crack, damage
model tables, bond types, densities, boundary conditions) runs once on the
host; every time step runs in the integrator, on the GPU.

Differences from the reference, all outside the time-stepping path:
* meshes are read by :mod:`.mesh` (legacy VTK, gmsh 2.2) instead of meshio,
  and ``mesh_file`` may also be an (nnodes, 3) array of coordinates;
* the neighbour list comes from libpd_hip's cell-list search
  (``pd_neighbour_list``): the same neighbour sets as the reference's
  KD-tree query (distance <= horizon), rows sorted ascending.  The
  reference's rows come in its KD-tree's internal order, which its fixtures
  (peripy/test/data/expected_connectivity*.npz) record and which today's
  scikit-learn no longer reproduces either; everything downstream is
  defined per bond, so only the slot order differs;
* model arrays given ``write_path`` are saved as .npy files (h5py is absent);
* the stiffness-correction factors (surface_correction, volume_correction,
  micromodulus_function) come from :mod:`.correction`, on the integrator's
  GPU when it has one.
"""
import pathlib
import warnings
from collections import namedtuple

import numpy as np

from . import _lib, crack as _crack, mesh as _mesh
from .correction import (set_imprecise_surface_correction,
                         set_micromodulus_function,
                         set_precise_surface_correction,
                         set_volume_correction)
from .integrators import Integrator

_MeshElements = namedtuple("MeshElements", ["connectivity", "boundary"])
_mesh_elements_2d = _MeshElements(connectivity="triangle", boundary="line")
_mesh_elements_3d = _MeshElements(connectivity="tetra", boundary="triangle")


def write_array(write_path, dataset, array):
    """Save a model array as ``write_path/<dataset>.npy``."""
    path = pathlib.Path(write_path)
    path.mkdir(parents=True, exist_ok=True)
    np.save(path / f"{dataset}.npy", array)


def read_array(read_path, dataset):
    """Load a model array saved by :func:`write_array` (or None)."""
    f = pathlib.Path(read_path) / f"{dataset}.npy"
    return np.load(f) if f.exists() else None


def vectorised(fn):
    """Mark a set-up callback as one that works on many nodes at once.

    The reference calls ``is_bond_type(x, y)`` once per bond and
    ``is_density(x)``, ``is_displacement_boundary(x)``,
    ``is_force_boundary(x)`` and ``is_tip(x)`` once per node from Python
    (model.py:657-780, 1093-1212): 1e8 calls on a million-node mesh.  A
    function marked with this decorator is called instead with coordinate
    arrays of shape (3, M) -- component first, so ``x[0]`` is the x
    coordinate of every point, as for a single node -- and must return the
    results for all M points at once:

    * ``is_bond_type``: M integer bond types.  Bonds are passed in the
      reference's order (node i ascending, then i's row order) in blocks of
      at most 4M bonds; the range checks are the reference's.
    * ``is_density``: M float densities.
    * ``is_displacement_boundary``, ``is_force_boundary``: a float array of
      shape (3, M), the value of DOF d of point m at [d, m], NaN where the
      per-node function would return None.
    * ``is_tip``: an array of shape (3, M) (any array-like; it is read as
      objects) holding what the per-node function would return per DOF:
      None, a tip name, or a tuple of names.

    The arrays built from them equal those of the per-node calls.
    """
    fn.vectorised = True
    return fn


def _is_vectorised(fn):
    return getattr(fn, "vectorised", False) is True


@vectorised
def _no_condition(x):
    """No boundary condition anywhere (the default callbacks)."""
    return np.full((3, np.shape(x)[1]), np.nan)


@vectorised
def _no_tip(x):
    """No tip anywhere."""
    return np.full((3, np.shape(x)[1]), None, dtype=object)


def create_crack(crack, nlist, n_neigh):
    """Remove bonds (create_crack.pyx:1-43), in place: for each pair (i, j),
    j leaves row i by swapping in the last live entry, the vacated last slot
    is set to -1 and n_neigh[i] decreases; then the same for i in row j.
    Runs in libpd_hip (``pd_create_crack``); nlist and n_neigh are int32."""
    _crack.create_crack(crack, nlist, n_neigh)


class Model(object):
    """A bond-based peridynamics model (peripy/model.py:24-1575)."""

    def __init__(self, mesh_file, integrator, horizon, critical_stretch,
                 bond_stiffness, transfinite=0,
                 volume_total=None, write_path=None, connectivity=None,
                 family=None, volume=None, initial_crack=None, dimensions=2,
                 is_density=None, is_bond_type=None,
                 is_displacement_boundary=None, is_force_boundary=None,
                 is_tip=None, density=None, bond_types=None,
                 stiffness_corrections=None,
                 surface_correction=None, volume_correction=None,
                 micromodulus_function=None, node_radius=None):
        """Create a :class:`Model` (model.py:42-479)."""
        if not isinstance(integrator, Integrator):
            raise InvalidIntegrator(integrator)
        self.integrator = integrator
        self.write_path = (None if write_path is None
                           else pathlib.Path(write_path))
        self.dimensions = dimensions
        if dimensions == 2:
            self.mesh_elements = _mesh_elements_2d
        elif dimensions == 3:
            self.mesh_elements = _mesh_elements_3d
        else:
            raise DimensionalityError(dimensions)

        self._read_mesh(mesh_file, transfinite)

        if volume is None:
            this_may_take_a_while(self.nnodes, "volume")
            self.volume = self._volume(transfinite, volume_total)
            if self.write_path is not None:
                write_array(self.write_path, "volume", self.volume)
        elif type(volume) == np.ndarray:
            if np.shape(volume) != (self.nnodes, ):
                raise ValueError("volume shape is wrong, and must be "
                                 "(nnodes, ) (expected {}, got {})".format(
                                     (self.nnodes, ), np.shape(volume)))
            warnings.warn("Reading volume from argument.")
            self.volume = volume.astype(np.float64)
        else:
            raise TypeError("volume type is wrong (expected {}, got "
                            "{})".format(type(volume), np.ndarray))

        if volume_correction is not None:
            if not ((type(node_radius) == float)
                    or (type(node_radius) == np.float64)):
                suggested = 0.5 * np.power(np.sum(self.volume) / self.nnodes,
                                           1. / 3)
                raise TypeError(
                    "If volume_correction (= {}) is applied, an average "
                    "node radius must be supplied as the keyword argument "
                    "node_radius, a float (got {}; half the cube root of "
                    "the mean node volume is {})".format(
                        volume_correction, type(node_radius), suggested))
            # partial volumes up to a node radius beyond the horizon take
            # part in the force integral, so the search reaches that far
            self.horizon = horizon + node_radius
        else:
            self.horizon = horizon

        if family is None or connectivity is None:
            this_may_take_a_while(self.nnodes, "family, connectivity")
            (self.family, nlist, n_neigh,
             self.max_neighbours) = self._set_neighbour_list(
                self.coords, self.horizon, self.nnodes, initial_crack,
                integrator.context)
            if self.write_path is not None:
                write_array(self.write_path, "family", self.family)
                write_array(self.write_path, "nlist", nlist)
                write_array(self.write_path, "n_neigh", n_neigh)
        else:
            if type(family) != np.ndarray:
                raise TypeError("family type is wrong (expected {}, got "
                                "{})".format(type(family), np.ndarray))
            if np.shape(family) != (self.nnodes, ):
                raise ValueError("family shape is wrong, and must be "
                                 "(nnodes, ) (expected {}, got {})".format(
                                     (self.nnodes, ), np.shape(family)))
            warnings.warn("Reading family from argument.")
            self.family = family.astype(np.intc)
            if type(connectivity) != tuple:
                raise TypeError("connectivity type is wrong (expected {} or"
                                " {}, got {})".format(
                                    tuple, type(None), type(connectivity)))
            if len(connectivity) != 2:
                raise ValueError("connectivity size is wrong (expected 2,"
                                 " got {})".format(len(connectivity)))
            warnings.warn("Reading connectivity from argument.")
            nlist, n_neigh = connectivity
            nlist = nlist.astype(np.intc)
            n_neigh = n_neigh.astype(np.intc)
            self.max_neighbours = np.intc(np.shape(nlist)[1])
            if integrator.context is None:
                if self.max_neighbours != self.family.max():
                    raise ValueError(
                        "max_neighbours, which is equal to the"
                        " size of axis 1 of nlist is wrong (expected "
                        " max_neighbours = np.shape(nlist)[1] = "
                        "family.max() = {}, got {})".format(
                            self.family.max(), self.max_neighbours))
            elif self.max_neighbours & (self.max_neighbours - 1):
                raise ValueError(
                    "max_neighbours, which is equal to the"
                    " size of axis 1 of nlist is wrong (expected "
                    " max_neighbours = np.shape(nlist)[1] = {},"
                    " got {})".format(
                        1 << (int(self.family.max() - 1)).bit_length(),
                        self.max_neighbours))

        if np.any(self.family == 0):
            raise FamilyError(self.family)

        self.initial_connectivity = (nlist, n_neigh)
        self.degrees_freedom = 3

        if stiffness_corrections is None:
            # model.py:349-383: ones, times the micromodulus values, the
            # volume corrections and the surface corrections, in that order
            applied = False
            stiffness_corrections = np.ones(
                (self.nnodes, self.max_neighbours), dtype=np.float64)
            if micromodulus_function is not None:
                stiffness_corrections = self._set_micromodulus_values(
                    micromodulus_function, stiffness_corrections, horizon)
                applied = True
            if volume_correction is not None:
                # model.py:363-365 passes node_radius in the horizon position
                # and horizon in the node_radius position; kept, so that the
                # factors are the reference's
                stiffness_corrections = self._set_volume_corrections(
                    volume_correction, stiffness_corrections, node_radius,
                    horizon)
                applied = True
            if surface_correction is not None:
                stiffness_corrections = self._set_surface_corrections(
                    surface_correction, stiffness_corrections)
                applied = True
            if applied:
                self.stiffness_corrections = stiffness_corrections
                if self.write_path is not None:
                    write_array(self.write_path, "stiffness_corrections",
                                stiffness_corrections)
            else:
                self.stiffness_corrections = None
        elif type(stiffness_corrections) == np.ndarray:
            if np.shape(stiffness_corrections) != (self.nnodes,
                                                   self.max_neighbours):
                raise ValueError("stiffness_corrections shape is wrong, "
                                 "and must be (nnodes, max_neighbours) "
                                 "(expected {}, got {})".format(
                                     (self.nnodes, self.max_neighbours),
                                     np.shape(stiffness_corrections)))
            warnings.warn(
                "Reading stiffness_corrections from argument; the flags "
                "surface_correction={}, volume_correction={} and "
                "micromodulus_function={} are not applied.".format(
                    surface_correction, volume_correction,
                    micromodulus_function))
            self.stiffness_corrections = stiffness_corrections.astype(
                np.float64)
        else:
            raise TypeError("stiffness_corrections type is wrong (expected {}"
                            ", got {})".format(np.ndarray,
                                               type(stiffness_corrections)))

        if is_bond_type is None:
            def is_bond_type(x, y):
                return 0

        (self.bond_stiffness, self.critical_stretch, self.plus_cs,
         self.nbond_types, self.nregimes) = self._set_damage_model(
            bond_stiffness, critical_stretch)

        if bond_types is None:
            self.bond_types = self._set_bond_types(
                self.initial_connectivity, is_bond_type, self.nbond_types,
                self.nregimes)
        elif type(bond_types) == np.ndarray:
            if np.shape(bond_types) != (self.nnodes, self.max_neighbours):
                raise ValueError("bond_types shape is wrong, "
                                 "and must be (nnodes, max_neighbours) "
                                 "(expected {}, got {})".format(
                                     (self.nnodes, self.max_neighbours),
                                     np.shape(bond_types)))
            warnings.warn("Reading bond_types from argument.")
            self.bond_types = bond_types.astype(np.intc)
        else:
            raise TypeError("bond_types type is wrong (expected {}"
                            ", got {})".format(np.ndarray, type(bond_types)))

        self.densities = self._set_densities(density, is_density)

        if is_force_boundary is None:
            is_force_boundary = _no_condition
        if is_displacement_boundary is None:
            is_displacement_boundary = _no_condition
        if is_tip is None:
            is_tip = _no_tip

        (self.bc_types, self.bc_values, self.force_bc_types,
         self.force_bc_values, self.tip_types,
         self.ntips) = self._set_boundary_conditions(
            is_displacement_boundary, is_force_boundary, is_tip)

        self.integrator.build(
            self.nnodes, self.degrees_freedom, self.max_neighbours,
            self.coords, self.volume, self.family, self.bc_types,
            self.bc_values, self.force_bc_types, self.force_bc_values,
            self.stiffness_corrections, self.bond_types, self.densities)

    # ------------------------------------------------------------------
    # set-up
    # ------------------------------------------------------------------
    def _read_mesh(self, filename, transfinite):
        """Coordinates and cells (model.py:481-503)."""
        if isinstance(filename, np.ndarray):
            self.coords = np.ascontiguousarray(filename, dtype=np.float64)
            if self.coords.ndim != 2 or self.coords.shape[1] != 3:
                raise ValueError("coordinates must be an (nnodes, 3) array")
            self.nnodes = self.coords.shape[0]
            self.mesh_connectivity = None
            self.mesh_boundary = None
            return
        points, cells = _mesh.read(filename)
        self.coords = np.ascontiguousarray(points, dtype=np.float64)
        self.nnodes = self.coords.shape[0]
        self.mesh_connectivity = None
        self.mesh_boundary = None
        if not transfinite:
            self.mesh_connectivity = cells[self.mesh_elements.connectivity]
            self.mesh_boundary = cells.get(self.mesh_elements.boundary)

    def write_mesh(self, filename, damage=None, displacements=None,
                   file_format=None):
        """Write nodes, cells, damage and displacements (model.py:505-534)."""
        cells = []
        if self.mesh_connectivity is not None:
            cells.append((self.mesh_elements.connectivity,
                          self.mesh_connectivity))
        if self.mesh_boundary is not None:
            cells.append((self.mesh_elements.boundary, self.mesh_boundary))
        _mesh.write_vtk(filename, self.coords, cells,
                        {"damage": damage, "displacements": displacements})

    def _setup_device(self):
        """GPU for set-up work: the integrator's, if present, else -1
        (the host)."""
        context = self.integrator.context
        device = (getattr(context, "device", 0) if context
                  else getattr(self.integrator, "device", 0))
        try:
            device = int(device)
        except (TypeError, ValueError):
            return -1
        return device if _lib.device_available(device) else -1

    def _set_micromodulus_values(self, micromodulus_function,
                                 stiffness_corrections, horizon):
        """Multiply in the conical micromodulus values (model.py:782-821);
        only micromodulus_function == 0 exists."""
        nlist, n_neigh = self.initial_connectivity
        if micromodulus_function == 0:
            set_micromodulus_function(
                stiffness_corrections, self.coords, nlist, n_neigh,
                np.float64(horizon), np.intc(micromodulus_function),
                device=self._setup_device())
        else:
            raise ValueError("micromodulus_function value is wrong "
                             "(expected 0 or None, got {})".format(
                                 micromodulus_function))
        return stiffness_corrections

    def _set_volume_corrections(self, volume_correction,
                                stiffness_corrections, horizon, node_radius):
        """Multiply in the partial-volume factors (model.py:823-876); only
        volume_correction == 0 (the HHB algorithm) exists."""
        nlist, n_neigh = self.initial_connectivity
        if volume_correction == 0:
            set_volume_correction(
                stiffness_corrections, self.coords, nlist, n_neigh,
                np.float64(horizon), np.float64(node_radius),
                np.intc(volume_correction), device=self._setup_device())
        else:
            raise ValueError("volume_correction value is wrong "
                             "(expected 0 or None, got {})".format(
                                 volume_correction))
        return stiffness_corrections

    def _set_surface_corrections(self, surface_correction,
                                 stiffness_corrections):
        """Multiply in the volume-method surface factors (model.py:878-939):
        1 uses the nodal volumes, 0 their mean.  The bulk family volume is
        pi h^2 (2-D) or 4/3 pi h^3 (3-D) of the model's horizon."""
        nlist, n_neigh = self.initial_connectivity
        if self.dimensions == 2:
            family_volume_bulk = np.float64(np.pi * np.power(self.horizon, 2))
        else:
            family_volume_bulk = np.float64(
                (4. / 3) * np.pi * np.power(self.horizon, 3))
        if surface_correction == 1:
            set_precise_surface_correction(
                stiffness_corrections, nlist, n_neigh, self.volume,
                family_volume_bulk, device=self._setup_device())
        elif surface_correction == 0:
            average_node_volume = np.float64(np.sum(self.volume) / self.nnodes)
            set_imprecise_surface_correction(
                stiffness_corrections, nlist, n_neigh, average_node_volume,
                family_volume_bulk, device=self._setup_device())
        else:
            raise ValueError("surface_correction value is wrong "
                             "(expected 0, 1 or None, got {})".format(
                                 surface_correction))
        return stiffness_corrections

    def _set_neighbour_list(self, coords, horizon, nnodes,
                            initial_crack=None, context=None):
        """family, nlist, n_neigh, max_neighbours (model.py:536-597).

        Layout as model.py:575-582: a truthy context gives a power-of-two
        width padded with -1, otherwise width family.max() padded with 0.
        With a HIP context the search runs on its GPU (same result as the
        host search).
        """
        device = getattr(context, "device", None) if context else None
        if device is not None and not _lib.device_available(device):
            device = None
        family = _lib.family_count(coords, horizon, device)
        if context:
            max_neighbours = np.intc(1 << (int(family.max() - 1)).bit_length())
            pad = -1
        else:
            max_neighbours = np.intc(family.max())
            pad = 0
        family, nlist = _lib.neighbour_list(coords, horizon,
                                            int(max_neighbours), pad, device)
        n_neigh = family.copy()
        if initial_crack is not None:
            if callable(initial_crack):
                initial_crack = initial_crack(coords, nlist, n_neigh)
            create_crack(np.array(initial_crack, dtype=np.int32), nlist,
                         n_neigh)
        return family, nlist, n_neigh, max_neighbours

    def _volume(self, transfinite, volume_total):
        """Node volumes (model.py:599-655)."""
        if transfinite:
            if volume_total is None:
                raise TypeError("In transfinite mode, a total mesh volume "
                                "volume_total' must be provided as a keyword"
                                " argument (expected {}, got {})".format(
                                    float, type(volume_total)))
            return (volume_total / self.nnodes) * np.ones(self.nnodes)
        if self.mesh_connectivity is None:
            raise TypeError("a volume array or transfinite mode is needed "
                            "when the mesh has no cells")
        volume = np.zeros(self.nnodes)
        cells = self.mesh_connectivity
        c = self.coords
        if self.dimensions == 2:
            a, b, d = c[cells[:, 0]], c[cells[:, 1]], c[cells[:, 2]]
            ev = 0.5 * np.linalg.norm(np.cross(b - a, d - a), axis=1)
            per = 3
        else:
            a, b, cc, d = (c[cells[:, 0]], c[cells[:, 1]], c[cells[:, 2]],
                           c[cells[:, 3]])
            ev = np.abs(np.einsum("ij,ij->i", a - d,
                                  np.cross(b - d, cc - d))) / 6
            per = 4
        # element by element, in order, as the reference accumulates
        for q in range(cells.shape[0]):
            volume[cells[q]] += ev[q] / per
        return volume.astype(np.float64)

    def _set_densities(self, density, is_density):
        """(nnodes, 3) densities or None (model.py:657-707)."""
        if density is None:
            if is_density is None:
                return None
            if not callable(is_density):
                raise TypeError("is_density must be a *function*.")
            if _is_vectorised(is_density):
                density = np.asarray(is_density(self.coords.T.copy()))
                if density.shape != (self.nnodes,) or density.dtype.kind != "f":
                    raise TypeError(
                        "a vectorised is_density must return a float array of "
                        "shape (nnodes,) (got {} of shape {})".format(
                            density.dtype, density.shape))
                density = density.astype(np.float64)
                if self.write_path is not None:
                    write_array(self.write_path, "density", density)
                return np.transpose(np.tile(density, (self.degrees_freedom, 1))
                                    ).astype(np.float64)
            if type(is_density(self.coords[0])) is not float:
                raise TypeError(
                    "is_density must be a function that returns a *float* "
                    "(expected {}, got {})".format(
                        float, type(is_density(self.coords[0]))))
            density = np.array([is_density(x) for x in self.coords],
                               dtype=np.float64)
            if self.write_path is not None:
                write_array(self.write_path, "density", density)
        elif type(density) == np.ndarray:
            if np.shape(density) != (self.nnodes,):
                raise ValueError("densty shape is wrong, and must be "
                                 "(nnodes,) (expected {}, got {})".format(
                                     (self.nnodes,), np.shape(density)))
            warnings.warn("Reading density from argument.")
        else:
            raise TypeError("density type is wrong, and must be an array of"
                            " shape (nnodes,) (expected {}, got {})".format(
                                np.ndarray, type(density)))
        return np.transpose(np.tile(density, (self.degrees_freedom, 1))
                            ).astype(np.float64)

    def _set_bond_types(self, connectivity, is_bond_type, nbond_types,
                        nregimes):
        """(nnodes, max_neighbours) bond types or None (model.py:709-780)."""
        if nbond_types != 1:
            if not callable(is_bond_type):
                raise TypeError("is_bond_type must be a *function*.")
            nlist, n_neigh = connectivity
            bond_types = np.zeros((self.nnodes, self.max_neighbours),
                                  dtype=np.intc)
            if callable(getattr(is_bond_type, "bond_types", None)):
                # a built-in predicate (predicates.PlaneBondTypes): every
                # bond at once in libpd_hip, with the reference's checks
                bond_types = is_bond_type.bond_types(self.coords, nlist,
                                                     n_neigh)
                live = (np.arange(self.max_neighbours)[None, :]
                        < np.asarray(n_neigh)[:, None])
                self._check_bond_types(bond_types[live], nbond_types)
            elif _is_vectorised(is_bond_type):
                self._bond_types_vectorised(is_bond_type, nlist, n_neigh,
                                            nbond_types, bond_types)
            if (_is_vectorised(is_bond_type)
                    or callable(getattr(is_bond_type, "bond_types", None))):
                if self.write_path is not None:
                    write_array(self.write_path, "bond_types", bond_types)
                return bond_types
            for i in range(self.nnodes):
                for neigh in range(n_neigh[i]):
                    j = nlist[i][neigh]
                    bt = is_bond_type(self.coords[i, :], self.coords[j, :])
                    if type(bt) is not int:
                        raise TypeError(
                            "is_bond_type must be a function that returns "
                            "an *int* (expected {}, got {})".format(
                                int, type(bt)))
                    if bt < 0:
                        raise ValueError(
                            "is_bond_type must be a function that returns "
                            "a *positive* int or 0 (got {})".format(bt))
                    if bt > nbond_types - 1:
                        raise ValueError(
                            "is_bond_type must be a function that returns "
                            "a positive int or 0 which is *less* than "
                            "nbond_types (nbond_types = {}, got {})".format(
                                nbond_types, bt))
                    bond_types[i][neigh] = bt
        elif nregimes != 1:
            bond_types = np.zeros((self.nnodes, self.max_neighbours),
                                  dtype=np.intc)
        else:
            return None
        if self.write_path is not None:
            write_array(self.write_path, "bond_types", bond_types)
        return bond_types

    def _bond_types_vectorised(self, is_bond_type, nlist, n_neigh,
                               nbond_types, out, block=1 << 22):
        """Bond types from a :func:`vectorised` is_bond_type, written into
        ``out`` row by row in blocks of at most ``block`` bonds."""
        nlist = np.asarray(nlist)
        n_neigh = np.asarray(n_neigh)
        cum = np.concatenate([[0], np.cumsum(n_neigh, dtype=np.int64)])
        lo = 0
        while lo < self.nnodes:
            hi = int(np.searchsorted(cum, cum[lo] + block, side="right")) - 1
            hi = min(max(hi, lo + 1), self.nnodes)
            rows = nlist[lo:hi]
            live = np.arange(self.max_neighbours)[None, :] < n_neigh[lo:hi,
                                                                     None]
            i = np.repeat(np.arange(lo, hi), n_neigh[lo:hi])
            j = rows[live]
            # component-major views: x[0] is every bond's x coordinate
            bt = np.asarray(is_bond_type(self.coords[i].T,
                                         self.coords[j].T))
            if bt.shape != (len(i),) or bt.dtype.kind not in "iu":
                raise TypeError(
                    "a vectorised is_bond_type must return an integer array "
                    "with one entry per bond (got {} of shape {} for {} "
                    "bonds)".format(bt.dtype, bt.shape, len(i)))
            self._check_bond_types(bt, nbond_types)
            sub = out[lo:hi]
            sub[live] = bt
            lo = hi

    @staticmethod
    def _check_bond_types(bt, nbond_types):
        """The range checks of model.py:741-756 on many bond types."""
        bad = np.nonzero((bt < 0) | (bt > nbond_types - 1))[0]
        if bad.size:
            q = bad[0]
            if bt[q] < 0:
                raise ValueError(
                    "is_bond_type must be a function that returns "
                    "a *positive* int or 0 (got {})".format(bt[q]))
            raise ValueError(
                "is_bond_type must be a function that returns "
                "a positive int or 0 which is *less* than "
                "nbond_types (nbond_types = {}, got {})".format(
                    nbond_types, bt[q]))

    def _set_damage_model(self, bond_stiffness, critical_stretch):
        """Damage-model tables and plus_cs (model.py:941-1091).

        Tables are (nbond_types, nregimes); plus_cs holds the intercepts
        c_r = c_{r-1} + k_{r-1} s_{r-1} - k_r s_{r-1} (c_0 = 0) that keep the
        n-linear force continuous at each breakpoint.
        """
        if type(bond_stiffness) != type(critical_stretch):
            raise TypeError(
                "bond_stiffness must be the same type "
                "as critical_stretch (expected {}, got {})".format(
                    type(critical_stretch), type(bond_stiffness)))
        if type(bond_stiffness) in (list, np.ndarray):
            if np.shape(bond_stiffness) != np.shape(critical_stretch):
                raise ValueError(
                    "The shape of bond_stiffness must be equal to the shape "
                    "of critical_stretch (expected {}, got {})".format(
                        np.shape(critical_stretch),
                        np.shape(bond_stiffness)))
            if np.shape(bond_stiffness) == (1,):
                nregimes = nbond_types = 1
                bond_stiffness = np.float64(bond_stiffness[0])
                critical_stretch = np.float64(critical_stretch[0])
                plus_cs = None
            elif np.shape(bond_stiffness) == ():
                nregimes = nbond_types = 1
                bond_stiffness = np.float64(bond_stiffness)
                critical_stretch = np.array(critical_stretch)
                plus_cs = None
            elif np.shape(bond_stiffness[0]) == (1,):
                nregimes = 1
                nbond_types = np.shape(bond_stiffness)[0]
                bond_stiffness = np.array(bond_stiffness, dtype=np.float64)
                critical_stretch = np.array(critical_stretch,
                                            dtype=np.float64)
                plus_cs = np.zeros(nbond_types).astype(np.float64)
            elif np.shape(bond_stiffness[0]) == ():
                nbond_types = 1
                nregimes = np.shape(bond_stiffness)[0]
                if not all(critical_stretch[i] <= critical_stretch[i + 1]
                           for i in range(nregimes - 1)):
                    raise DamageModelError(critical_stretch)
                bond_stiffness = np.array(bond_stiffness, dtype=np.float64)
                critical_stretch = np.array(critical_stretch,
                                            dtype=np.float64)
                plus_cs = np.zeros(nregimes)
                c_prev = 0.0
                for r in range(1, nregimes):
                    c_i = (c_prev
                           + (bond_stiffness[r - 1] * critical_stretch[r - 1])
                           - (bond_stiffness[r] * critical_stretch[r - 1]))
                    plus_cs[r] = c_i
                    c_prev = c_i
                plus_cs = plus_cs.astype(np.float64)
            else:
                nregimes = np.shape(bond_stiffness)[1]
                nbond_types = np.shape(bond_stiffness)[0]
                for i in range(nbond_types):
                    if not all(critical_stretch[i][j] <=
                               critical_stretch[i][j + 1]
                               for j in range(nregimes - 1)):
                        raise DamageModelError(critical_stretch[i])
                bond_stiffness = np.array(bond_stiffness, dtype=np.float64)
                critical_stretch = np.array(critical_stretch,
                                            dtype=np.float64)
                plus_cs = np.zeros((nbond_types, nregimes))
                c_prev = np.zeros(nbond_types)
                for r in range(1, nregimes):
                    c_i = (c_prev
                           + (bond_stiffness[:, r - 1]
                              * critical_stretch[:, r - 1])
                           - (bond_stiffness[:, r]
                              * critical_stretch[:, r - 1]))
                    plus_cs[:, r] = c_i
                    c_prev = c_i
                plus_cs = plus_cs.astype(np.float64)
        elif type(bond_stiffness) in (float, np.float64):
            nregimes = nbond_types = 1
            bond_stiffness = np.float64(bond_stiffness)
            critical_stretch = np.float64(critical_stretch)
            plus_cs = None
        else:
            raise TypeError(
                "Type of bond_stiffness and critical_stretch is not "
                "supported (expected {} or {}, got {})".format(
                    float, np.ndarray, type(bond_stiffness)))
        nregimes = np.intc(nregimes)
        nbond_types = np.intc(nbond_types)
        if np.any(critical_stretch < 0):
            raise ValueError("critical_stretch values must not be < 0, "
                             "(got {})".format(critical_stretch))
        return (bond_stiffness, critical_stretch, plus_cs, nbond_types,
                nregimes)

    def _set_boundary_conditions(self, is_displacement_boundary,
                                 is_force_boundary, is_tip):
        """BC and tip arrays (model.py:1093-1212).

        force_bc_values are force densities: the value divided by the node's
        volume and by the number of force-BC nodes (model.py:1194-1209).
        """
        functions = {"is_displacement_boundary": is_displacement_boundary,
                     "is_force_boundary": is_force_boundary,
                     "is_tip": is_tip}
        for name, fn in functions.items():
            if not callable(fn):
                raise TypeError("{} must be a *function*.".format(name))
            if _is_vectorised(fn):
                continue
            if type(fn(self.coords[0][:])) is not list:
                raise TypeError(
                    "{} must be a function that returns a *list*.".format(
                        name))
            if len(fn(self.coords[0][:])) != 3:
                raise TypeError("{} must be a function that returns a list"
                                " of length *3*".format(name))

        bc_types = np.zeros((self.nnodes, 3), dtype=np.intc)
        bc_values = np.zeros((self.nnodes, 3), dtype=np.float64)
        force_bc_types = np.zeros((self.nnodes, 3), dtype=np.intc)
        force_bc_values = np.zeros((self.nnodes, 3), dtype=np.float64)
        tip_types = {}
        ntips = {"model": self.nnodes}
        num_force_bc_nodes = 0

        def set_tip(tip, i, j):
            key = str(tip)
            if key not in ntips:
                ntips[key] = 0
                tip_types[key] = []
            ntips[key] += 1
            tip_types[key].append((i, j))

        # the vectorised callbacks, all nodes at once (NaN / None: no entry)
        xt = self.coords.T.copy()
        vec_bnd = vec_fbnd = vec_tip = None
        if _is_vectorised(is_displacement_boundary):
            vec_bnd = self._vector_conditions(is_displacement_boundary,
                                              "is_displacement_boundary", xt)
            set_ = ~np.isnan(vec_bnd)
            bc_types[set_] = 1
            bc_values[set_] = vec_bnd[set_]
        if _is_vectorised(is_force_boundary):
            vec_fbnd = self._vector_conditions(is_force_boundary,
                                               "is_force_boundary", xt)
            set_ = ~np.isnan(vec_fbnd)
            force_bc_types[set_] = 1
            force_bc_values[set_] = (vec_fbnd
                                     / self.volume[:, np.newaxis])[set_]
            num_force_bc_nodes = int(np.count_nonzero(set_.any(axis=1)))
        if _is_vectorised(is_tip):
            vec_tip = np.asarray(is_tip(xt), dtype=object)
            if vec_tip.shape != (3, self.nnodes):
                raise TypeError(
                    "a vectorised is_tip must return an array of shape "
                    "(3, nnodes) (got shape {})".format(vec_tip.shape))
            vec_tip = vec_tip.T
        per_node = vec_bnd is None or vec_fbnd is None or vec_tip is None
        for i in range(self.nnodes if per_node else 0):
            bnd = (None if vec_bnd is not None
                   else is_displacement_boundary(self.coords[i][:]))
            fbnd = (None if vec_fbnd is not None
                    else is_force_boundary(self.coords[i][:]))
            tip = None if vec_tip is not None else is_tip(self.coords[i][:])
            is_force_node = 0
            for j in range(3):
                if bnd is not None and bnd[j] is not None:
                    bc_types[i, j] = np.intc(1)
                    bc_values[i, j] = np.float64(bnd[j])
                if fbnd is not None and fbnd[j] is not None:
                    is_force_node = 1
                    force_bc_types[i, j] = np.intc(1)
                    force_bc_values[i, j] = np.float64(
                        fbnd[j] / self.volume[i])
                if tip is not None and tip[j] is not None:
                    if type(tip[j]) is tuple:
                        for t in tip[j]:
                            set_tip(t, i, j)
                    else:
                        set_tip(tip[j], i, j)
            num_force_bc_nodes += is_force_node
        if vec_tip is not None:
            # the per-node loop's order: node ascending, then DOF
            for i, j in zip(*np.nonzero(vec_tip != None)):  # noqa: E711
                t = vec_tip[i, j]
                for name in (t if type(t) is tuple else (t,)):
                    set_tip(name, int(i), int(j))
        if num_force_bc_nodes != 0:
            force_bc_values = np.float64(
                np.divide(force_bc_values, num_force_bc_nodes))
        return (bc_types, bc_values, force_bc_types, force_bc_values,
                tip_types, ntips)

    def _vector_conditions(self, fn, name, xt):
        """(nnodes, 3) values of a vectorised boundary-condition callback,
        NaN where there is no condition."""
        v = np.asarray(fn(xt))
        if v.shape != (3, self.nnodes) or v.dtype.kind != "f":
            raise TypeError(
                "a vectorised {} must return a float array of shape "
                "(3, nnodes), NaN for no condition (got {} of shape "
                "{})".format(name, v.dtype, v.shape))
        return v.T.astype(np.float64)

    # ------------------------------------------------------------------
    # simulation
    # ------------------------------------------------------------------
    def simulate(self, steps, u=None, ud=None, connectivity=None,
                 regimes=None, critical_stretch=None, bond_stiffness=None,
                 displacement_bc_magnitudes=None, force_bc_magnitudes=None,
                 first_step=1, write=None, write_path=None):
        """Run ``steps`` steps (model.py:1214-1385).

        Returns ``(u, damage, (nlist, n_neigh), force, ud, data)``.
        """
        (u, ud, udd, force, body_force, nlist, n_neigh,
         displacement_bc_magnitudes, force_bc_magnitudes, damage, data,
         nwrites, write_path) = self._simulate_initialise(
            steps, first_step, write, regimes, u, ud,
            displacement_bc_magnitudes, force_bc_magnitudes, connectivity,
            bond_stiffness, critical_stretch, write_path)

        # HIP integrators sum on the device and copy only u and damage (for
        # the mesh file) at a write; other integrators take the full write
        monitor = getattr(self.integrator, "monitor", None)
        if write and monitor is not None:
            self.integrator.set_tips(self.tip_types)
        # the steps between two writes go to the library's own step loop in
        # one call (the same launches as one call per step) unless the
        # integrator's step is overridden in Python
        run = getattr(self.integrator, "run", None)
        if (run is not None and type(self.integrator).__call__.__module__
                != Integrator.__module__):
            run = None
        last = first_step + steps - 1
        step = first_step
        while step <= last:
            if run is not None:
                end = last if not write else min(last,
                                                 -(-step // write) * write)
                run(end - step + 1, displacement_bc_magnitudes[step - 1:end],
                    force_bc_magnitudes[step - 1:end])
                step = end
            else:
                self.integrator(displacement_bc_magnitudes[step - 1],
                                force_bc_magnitudes[step - 1])
            if write and step % write == 0 and monitor is not None:
                u, damage = self.integrator.write_fields(u, damage)
                self.write_mesh(write_path / f"U_{step}.vtk", damage, u)
                ii = step // write - (first_step - 1) // write - 1
                sums, tip_sums = monitor()
                for tip_type in self.tip_types:
                    if tip_type not in data:
                        data[tip_type] = {
                            k: np.zeros(nwrites, dtype=np.float64)
                            for k in ("displacement", "velocity",
                                      "acceleration", "force", "body_force")}
                    for k, v in tip_sums[tip_type].items():
                        data[tip_type][k][ii] += v
                data["model"]["step"][ii] = step
                for k in ("displacement", "velocity", "acceleration", "force",
                          "body_force", "damage_sum"):
                    data["model"][k][ii] = sums[k]
                if sums["damage_sum"] > 0.05 * self.nnodes:
                    warnings.warn("Over 5% of bonds have broken! "
                                  "peridynamics simulation continuing")
            elif write and step % write == 0:
                (u, ud, udd, force, body_force, damage, nlist,
                 n_neigh) = self.integrator.write(
                    u, ud, udd, force, body_force, damage, nlist, n_neigh)
                self.write_mesh(write_path / f"U_{step}.vtk", damage, u)
                ii = step // write - (first_step - 1) // write - 1
                for tip_type, node_list in self.tip_types.items():
                    if tip_type not in data:
                        data[tip_type] = {
                            k: np.zeros(nwrites, dtype=np.float64)
                            for k in ("displacement", "velocity",
                                      "acceleration", "force", "body_force")}
                    for i, j in node_list:
                        data[tip_type]["displacement"][ii] += u[i, j]
                        data[tip_type]["velocity"][ii] += ud[i, j]
                        data[tip_type]["acceleration"][ii] += udd[i, j]
                        data[tip_type]["force"][ii] += (force[i, j]
                                                        * self.volume[i])
                        data[tip_type]["body_force"][ii] += (
                            body_force[i, j] * self.volume[i])
                data["model"]["step"][ii] = step
                data["model"]["displacement"][ii] = np.sum(u)
                data["model"]["velocity"][ii] = np.sum(ud)
                data["model"]["acceleration"][ii] = np.sum(udd)
                data["model"]["force"][ii] = np.sum(
                    force * self.volume[:, np.newaxis])
                data["model"]["body_force"][ii] = np.sum(
                    body_force * self.volume[:, np.newaxis])
                damage_sum = np.sum(damage)
                data["model"]["damage_sum"][ii] = damage_sum
                if damage_sum > 0.05 * self.nnodes:
                    warnings.warn("Over 5% of bonds have broken! "
                                  "peridynamics simulation continuing")
            step += 1
        for tip_type_str in data:
            ntip = self.ntips[tip_type_str]
            if ntip != 0:
                for k in ("displacement", "velocity", "acceleration"):
                    data[tip_type_str][k] /= ntip
        (u, ud, udd, force, body_force, damage, nlist,
         n_neigh) = self.integrator.write(u, ud, udd, force, body_force,
                                          damage, nlist, n_neigh)
        return (u, damage, (nlist, n_neigh), force, ud, data)

    def _simulate_initialise(self, steps, first_step, write, regimes, u, ud,
                             displacement_bc_magnitudes, force_bc_magnitudes,
                             connectivity, bond_stiffness, critical_stretch,
                             write_path):
        """Simulation state and buffers (model.py:1387-1575)."""
        if u is None:
            u = np.zeros((self.nnodes, 3), dtype=np.float64)
        if ud is None:
            ud = np.zeros((self.nnodes, 3), dtype=np.float64)
        force = np.zeros((self.nnodes, 3), dtype=np.float64)
        body_force = np.zeros((self.nnodes, 3), dtype=np.float64)
        damage = np.zeros(self.nnodes, dtype=np.float64)
        udd = np.zeros((self.nnodes, 3), dtype=np.float64)
        mags = []
        for name, m in (("displacement_bc_magnitudes",
                         displacement_bc_magnitudes),
                        ("force_bc_magnitudes", force_bc_magnitudes)):
            if m is None:
                m = np.zeros(first_step + steps - 1, dtype=np.float64)
            elif type(m) == np.ndarray:
                if len(m) < steps:
                    raise ValueError(
                        "{} length must be equal to or greater than "
                        "(first_step + steps - 1), (expected {}, got {})"
                        .format(name, first_step + steps - 1, len(m)))
                m = m.astype(np.float64)
            else:
                raise TypeError("{} type is wrong (expected {}, got {})"
                                .format(name, np.ndarray, type(m)))
            mags.append(m)
        displacement_bc_magnitudes, force_bc_magnitudes = mags
        if connectivity is None:
            nlist, n_neigh = self.initial_connectivity
            nlist = nlist.copy()
            n_neigh = n_neigh.copy()
        elif type(connectivity) == tuple:
            if len(connectivity) != 2:
                raise ValueError("connectivity size is wrong (expected 2,"
                                 " got {})".format(len(connectivity)))
            nlist, n_neigh = connectivity
        else:
            raise TypeError("connectivity type is wrong (expected {} or"
                            " {}, got {})".format(tuple, type(None),
                                                  type(connectivity)))
        if regimes is None:
            regimes = np.zeros((self.nnodes, self.max_neighbours),
                               dtype=np.intc)
        elif type(regimes) == np.ndarray:
            if np.shape(regimes) != (self.nnodes, self.max_neighbours):
                raise ValueError("regimes shape is wrong, and must be "
                                 "(nnodes, max_neighbours) "
                                 "(expected {}, got {})".format(
                                     (self.nnodes, self.max_neighbours),
                                     np.shape(regimes)))
            regimes = regimes.astype(np.intc)
        else:
            raise TypeError("regimes type is wrong (expected {} or {}, got "
                            "{})".format(np.ndarray, type(None),
                                         type(regimes)))
        if bond_stiffness is None and critical_stretch is None:
            bond_stiffness = self.bond_stiffness
            critical_stretch = self.critical_stretch
            plus_cs = self.plus_cs
            nbond_types = self.nbond_types
            nregimes = self.nregimes
        else:
            (bond_stiffness, critical_stretch, plus_cs, nbond_types,
             nregimes) = self._set_damage_model(bond_stiffness,
                                                critical_stretch)
            if nbond_types != self.nbond_types:
                raise ValueError(
                    "Number of bond types has unexpectedly changed from when "
                    " the model was constructed. Please reinstantiate "
                    ":class:`Model` with the new number of bond types "
                    "(expected {}, got {}).".format(self.nbond_types,
                                                    nbond_types))
        write_path = (pathlib.Path() if write_path is None
                      else pathlib.Path(write_path))
        data = {}
        nwrites = None
        if write:
            nwrites = ((first_step + steps - 1) // write
                       - (first_step - 1) // write)
            data["model"] = {
                "step": np.zeros(nwrites, dtype=int),
                **{k: np.zeros(nwrites, dtype=np.float64)
                   for k in ("displacement", "velocity", "acceleration",
                             "force", "body_force", "damage_sum")}}
        self.integrator.create_buffers(
            nlist, n_neigh, bond_stiffness, critical_stretch, plus_cs, u, ud,
            udd, force, body_force, damage, regimes, nregimes, nbond_types)
        return (u, ud, udd, force, body_force, nlist, n_neigh,
                displacement_bc_magnitudes, force_bc_magnitudes, damage, data,
                nwrites, write_path)


def initial_crack_helper(crack_function):
    """Decorator: pairwise crack predicate -> initial_crack function.

    model.py:1578-1612: the returned function lists every bonded pair
    (i, j), i < j, in the order i ascending, j in nlist order, for which
    ``crack_function(coords[i], coords[j])`` is truthy.
    """
    def initial_crack(coords, nlist, n_neigh):
        crack = []
        for i in range(nlist.shape[0]):
            for j in nlist[i][0:n_neigh[i]]:
                if i < j and crack_function(coords[i], coords[j]):
                    crack.append((i, int(j)))
        return crack
    return initial_crack


def this_may_take_a_while(nnodes, calculation):
    """Warn for large set-up calculations (model.py:1615-1630)."""
    if nnodes > 9000:
        warnings.warn(f"Calculating {calculation} for {nnodes} nodes... "
                      "this may take a while.")


class DimensionalityError(Exception):
    """An invalid dimensionality argument used to construct a model."""

    def __init__(self, dimensions):
        super().__init__("The number of dimensions must be 2 or 3,"
                         f" {dimensions} was given.")


class FamilyError(Exception):
    """One or more nodes have no bonds in the initial state."""

    def __init__(self, family):
        idx = " ".join(f"{i}" for i in np.where(family == 0)[0])
        super().__init__("The following nodes have no bonds in the initial "
                         f"state, {idx}.")


class DamageModelError(Exception):
    """An invalid critical stretch argument was used to construct a model."""

    def __init__(self, critical_stretch):
        super().__init__(
            "The critical_stretch list or array for a bond-type with "
            "multiple regimes must be in ascending order, "
            f" {critical_stretch} was given.")


class InvalidIntegrator(Exception):
    """An invalid integrator has been passed to `Model`."""

    def __init__(self, integrator):
        super().__init__(f"{integrator} is not an instance of "
                         "probabilistic_peridynamics_amd.integrators."
                         "Integrator")
