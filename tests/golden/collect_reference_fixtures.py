"""Copy the reference's own test data files used as golden fixtures.

Run once, in the build container, where the reference tree is mounted.  It
copies DATA files only (numpy arrays and meshes that PeriPy's test suite
loads: peripy/test/data/*), never source.  The copies are committed next to
this script so that the tests run without the reference present.

Each fixture and the reference test that reads it:

  example_mesh.vtk               conftest.py:72 (2113-node 2-D plate)
  example_mesh_3d.vtk            test_model.py:56-88 (1175-node 3-D block)
  force_test.vtk                 test_model.py:1569 (3-node force sign model)
  expected_volume.npy            test_model.py:263
  expected_connectivity_basic.npz      test_model.py:317 (pre-crack, Cython)
  expected_connectivity_crack.npz      test_regression.py:62, test_integrators.py:177
  expected_connectivity_crack_cl.npz   test_regression.py:124, test_integrators.py:451
  expected_displacements.npy, expected_force.npy, expected_damage.npy
                                 test_regression.py:45-79, test_integrators.py:174-176
  expected_*_euler_cromer.npy    test_integrators.py:575-577 (+ unreferenced velocities)
  expected_*_velocity_verlet.npy test_integrators.py:721-727
  expected_bc_types/values.npy, expected_force_bc_types/values.npy
                                 test_model.py:1466-1471
  expected_corrections.npz       the twelve expected_micromodulus_values_*,
                                 expected_volume_corrections_* and
                                 expected_stiffness_corrections_* arrays of
                                 test_model.py:805-1029, packed unchanged (one
                                 key per original file stem) into one
                                 compressed archive
  expected_bond_types.npz        expected_bond_types(_cl).npy of
                                 test_model.py:543-561, packed unchanged
"""
import pathlib
import shutil
import sys

SRC = pathlib.Path("/root/reference/peripy/test/data")
DST = pathlib.Path(__file__).resolve().parent

FILES = [
    "example_mesh.vtk",
    "example_mesh_3d.vtk",
    "force_test.vtk",
    "expected_volume.npy",
    "expected_connectivity_basic.npz",
    "expected_connectivity_crack.npz",
    "expected_connectivity_crack_cl.npz",
    "expected_displacements.npy",
    "expected_force.npy",
    "expected_damage.npy",
    "expected_displacements_euler_cromer.npy",
    "expected_force_euler_cromer.npy",
    "expected_damage_euler_cromer.npy",
    "expected_velocities_euler_cromer.npy",
    "expected_displacements_velocity_verlet.npy",
    "expected_velocities_velocity_verlet.npy",
    "expected_accelerations_velocity_verlet.npy",
    "expected_force_velocity_verlet.npy",
    "expected_damage_velocity_verlet.npy",
    "expected_bc_types.npy",
    "expected_bc_values.npy",
    "expected_force_bc_types.npy",
    "expected_force_bc_values.npy",
]

PACKED = {
    "expected_corrections.npz": [
        "expected_micromodulus_values_2d",
        "expected_micromodulus_values_3d",
        "expected_volume_corrections_2d",
        "expected_volume_corrections_3d",
        "expected_stiffness_corrections_2d",
        "expected_stiffness_corrections_2d_cl",
        "expected_stiffness_corrections_2d_precise",
        "expected_stiffness_corrections_2d_precise_cl",
        "expected_stiffness_corrections_3d",
        "expected_stiffness_corrections_3d_cl",
        "expected_stiffness_corrections_3d_precise",
        "expected_stiffness_corrections_3d_precise_cl",
    ],
    "expected_bond_types.npz": [
        "expected_bond_types",
        "expected_bond_types_cl",
    ],
}

if __name__ == "__main__":
    if not SRC.is_dir():
        sys.exit(f"{SRC} not found: the fixtures are already committed.")
    for name in FILES:
        shutil.copyfile(SRC / name, DST / name)
        print("copied", name)
    import numpy as np
    for archive, stems in PACKED.items():
        np.savez_compressed(DST / archive, **{
            stem: np.load(SRC / f"{stem}.npy") for stem in stems})
        print("packed", archive)
