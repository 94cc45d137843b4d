"""The C-ABI libraries load and export every function their headers declare
(no compute calls: this runs without a GPU)."""
import ctypes
import pathlib
import re

import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent
HEADERS = {
    "tfs_solver.h": ROOT / "t-flows_amd" / "libtfs_solver.so",
    "tfs_mesh.h": ROOT / "t-flows_amd" / "libtfs_mesh.so",
}


def declared(header):
    text = (ROOT / "include" / header).read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set(re.findall(r"^[A-Za-z_][\w\s\*]*?\b(tfs_\w+)\s*\(", text, flags=re.M))
    return sorted(names)


@pytest.mark.parametrize("header", sorted(HEADERS))
def test_library_exports_declared_symbols(header):
    lib_path = HEADERS[header]
    if not lib_path.exists():
        pytest.fail(f"{lib_path.name} not built")
    lib = ctypes.CDLL(str(lib_path))
    names = declared(header)
    assert len(names) >= 5
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_solver_header_cites_reference():
    text = (ROOT / "include" / "tfs_solver.h").read_text()
    for f in ("Cg.f90", "Bicg.f90", "Cgs.f90", "Prec_Form.f90", "Prec_Solve.f90",
              "Residual.f90", "Matrix_Mod.f90"):
        assert f in text


def test_status_strings_without_gpu():
    lib = ctypes.CDLL(str(HEADERS["tfs_solver.h"]))
    lib.tfs_status_string.restype = ctypes.c_char_p
    lib.tfs_status_string.argtypes = [ctypes.c_int]
    assert lib.tfs_status_string(0) == b"ok"
    assert b"not converged" in lib.tfs_status_string(1)
    assert b"breakdown" in lib.tfs_status_string(-3)


def test_invalid_structure_rejected_without_gpu():
    """tfs_matrix_create validates the CRS layout before it touches a device."""
    import numpy as np
    lib = ctypes.CDLL(str(HEADERS["tfs_solver.h"]))
    p = ctypes.c_void_p()

    def create(row, col, dia, base=0):
        r = np.ascontiguousarray(row, np.int32)
        c = np.ascontiguousarray(col, np.int32)
        d = np.ascontiguousarray(dia, np.int32)
        return lib.tfs_matrix_create(ctypes.c_int(len(r) - 1), ctypes.c_longlong(len(c)),
                                     r.ctypes.data_as(ctypes.c_void_p),
                                     c.ctypes.data_as(ctypes.c_void_p),
                                     d.ctypes.data_as(ctypes.c_void_p), ctypes.c_int(base),
                                     ctypes.byref(p))
    assert create([0, 2, 4], [1, 0, 0, 1], [1, 3]) == -1   # columns not ascending
    assert create([0, 2, 4], [0, 1, 0, 1], [0, 0]) == -1   # dia of row 1 not its diagonal
    assert create([0, 2, 4], [0, 2, 0, 1], [0, 3]) == -1   # column out of range
    assert create([1, 3, 5], [1, 2, 1, 2], [1, 3], base=1) == -1  # 1-based, row 2 dia wrong
    assert create([0, 1, 1], [0], [0, 0]) == -1             # empty row (no diagonal)


def test_server_arguments_rejected_without_gpu():
    """tfs_server_create checks its arguments before it looks for a device."""
    lib = ctypes.CDLL(str(HEADERS["tfs_solver.h"]))
    p = ctypes.c_void_p()
    lib.tfs_server_create.argtypes = [ctypes.c_longlong, ctypes.c_int, ctypes.c_void_p]
    assert lib.tfs_server_create(0, 1, ctypes.byref(p)) == -1
    assert lib.tfs_server_create(1 << 30, 0, ctypes.byref(p)) == -1
    assert lib.tfs_server_create(1 << 30, 1, None) == -1
    lib.tfs_server_cancel.argtypes = [ctypes.c_void_p, ctypes.c_longlong]
    assert lib.tfs_server_cancel(None, 1) == -1


def test_matrix_info_layout_matches_the_header(tmp_path):
    """The ctypes mirror of tfs_matrix_info has the header's field names,
    offsets and size (a field appended on one side only would shift every
    later read)."""
    import shutil
    import subprocess
    from tflows_amd import solvers
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("no C compiler")
    info = solvers._Info
    names = [f[0] for f in info._fields_]
    src = tmp_path / "layout.c"
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "tfs_solver.h"',
             'int main(void) {', '  printf("size %zu\\n", sizeof(tfs_matrix_info));']
    lines += [f'  printf("{n} %zu\\n", offsetof(tfs_matrix_info, {n}));' for n in names]
    lines += ['  return 0;', '}']
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run([gcc, "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    out = dict(line.split() for line in subprocess.run(
        [str(exe)], check=True, capture_output=True, text=True).stdout.splitlines())
    assert int(out["size"]) == ctypes.sizeof(info)
    for n in names:
        assert int(out[n]) == getattr(info, n).offset, n
