"""Build libpd_hip.so in-tree with hipcc for gfx950.

``python -m probabilistic_peridynamics_amd.build`` compiles
``csrc/pd_hip.hip`` (time stepping), ``csrc/pd_nlist.hip`` (neighbour list),
``csrc/pd_corrections.hip`` (stiffness corrections), ``csrc/pd_crack.hip``
(initial crack) and ``csrc/pd_setup.hip`` (set-up predicates) into
``probabilistic_peridynamics_amd/libpd_hip.so``.  It
needs hipcc but no GPU.

Flags:
  -ffp-contract=off   no fused multiply-add unless written as fma(): the
                      stretch arithmetic must round exactly like the reference
                      (peripy/spatial.pyx compiled for baseline x86-64);
  -Xclang -target-feature -Xclang -packed-fp32-ops
                      keep packed FP32 instructions out of the generated code.
"""
import os
import pathlib
import shutil
import subprocess
import sys

HERE = pathlib.Path(__file__).resolve().parent
SOURCES = [HERE / "csrc" / name for name in (
    "pd_hip.hip", "pd_nlist.hip", "pd_corrections.hip", "pd_crack.hip",
    "pd_setup.hip")]
TARGET = HERE / "libpd_hip.so"
ARCH = os.environ.get("PD_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"),
                 "/opt/rocm/bin/hipcc"):
        if cand and pathlib.Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def build_library(verbose=False):
    """Compile libpd_hip.so; returns its path."""
    tmp = TARGET.with_suffix(".so.tmp")
    cmd = [hipcc(), *FLAGS, "-o", str(tmp), *map(str, SOURCES)]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building libpd_hip.so")
    os.replace(tmp, TARGET)
    return TARGET


if __name__ == "__main__":
    print(build_library(verbose=True))
