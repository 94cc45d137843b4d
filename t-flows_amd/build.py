"""In-tree build of the native libraries.

* libtfs_solver.so -- the HIP library (gfx950), hipcc;
* libtfs_mesh.so   -- host-side case generator and topology builder, g++.

Also builds the test-side checkers through oracle/Makefile: the C restatement
always, and the reference's own Fortran solver path when the reference tree is
present (it is not on the GPU machines, which use the prebuilt library).
"""
import os
import pathlib
import subprocess
import sys

PKG = pathlib.Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
INCLUDE = ROOT / "include"
REFERENCE_SOURCES = pathlib.Path("/root/reference/Sources")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off",
    # packed FP32 instructions stay off on this target
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]
HIP_SOURCES = ["tfs_solver.hip", "tfs_layout.cpp", "tfs_server.cpp", "tfs_assembly.hip",
               "tfs_topology.hip", "tfs_gradient.hip"]
HIP_DEPS = ["tfs_device.hpp", "tfs_kernels.hpp", "tfs_layout.hpp"]


def _run(cmd, cwd=None):
    print("+", " ".join(str(c) for c in cmd), flush=True)
    subprocess.run([str(c) for c in cmd], cwd=cwd, check=True)


def build_mesh():
    out = PKG / "libtfs_mesh.so"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-Wall",
          f"-I{INCLUDE}", f"-I{CSRC}", "-o", out, CSRC / "tfs_mesh.cpp", CSRC / "tfs_meshio.cpp", CSRC / "tfs_layout.cpp",
          CSRC / "tfs_layout_capi.cpp"])
    return out


def build_solver():
    out = PKG / "libtfs_solver.so"
    tmp = out.with_suffix(".so.tmp")
    _run([HIPCC, *HIP_FLAGS, f"-I{INCLUDE}", f"-I{CSRC}", "-o", tmp,
          *[CSRC / s for s in HIP_SOURCES]])
    os.replace(tmp, out)
    return out


def build_oracle():
    _run(["make", "-s", "-C", ROOT / "oracle"])
    if REFERENCE_SOURCES.is_dir():
        _run(["make", "-s", "-C", ROOT / "oracle", "ref"])
        # the Fortran drop-in linked with the reference's own Matrix_Mod
        # (needs libtfs_solver.so, built first by build_all)
        _run(["make", "-s", "-C", ROOT / "oracle", "shim"])


def build_all():
    build_mesh()
    build_solver()
    build_oracle()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"all": build_all, "mesh": build_mesh, "solver": build_solver,
         "oracle": build_oracle}[w]()
