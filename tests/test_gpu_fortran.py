"""The Fortran drop-in (t-flows_amd/fortran/solvers_mod.f90) called the way
T-FlowS calls it: a Fortran program whose Matrix_Type comes from the
reference's own Matrix_Mod_Topology (oracle/_ref/shim/shim_driver, built by
`make -C oracle shim` where the reference tree exists) calls Cg / Bicg / Cgs
of the drop-in for every preconditioner; its outputs are checked against the
oracle with the parity bar of the direct solves.  It then solves the same
system with the values assembled on the device (Solvers_Assemble_On_Device)
and handed over from a device copy (Solvers_Values_From_Device): both must
give the host-upload solve bit for bit."""
import json
import pathlib
import subprocess

import numpy as np
import pytest

from conftest import rel_l2
from tflows_amd import mesh

pytestmark = pytest.mark.gpu

ROOT = pathlib.Path(__file__).resolve().parent.parent
DRIVER = ROOT / "oracle" / "_ref" / "shim" / "shim_driver"


@pytest.mark.parametrize("kw", [
    dict(kind=mesh.MOMENTUM, nx=9, beta=2.0, jitter=0.1, permute=True, pin_first=False,
         rhs_zero_mean=False),
    dict(kind=mesh.MOMENTUM, nx=7, ny=11, nz=5, beta=1.0, jitter=0.05, permute=False,
         pin_first=False, rhs_zero_mean=False, seed=7),
], ids=["momentum9_permuted", "momentum_7x11x5"])
def test_fortran_drop_in(oracle, tmp_path, kw):
    if not DRIVER.exists():
        pytest.skip("Fortran driver not built (needs the reference's modules)")
    c = mesh.make_case(with_faces=True, **kw)
    nf = c.faces.size // 2
    with open(tmp_path / "faces.bin", "wb") as f:
        np.array([c.n, nf], np.int32).tofile(f)
        c.faces.astype(np.int32).tofile(f)
    a12, a21, dex = c.face_coefs
    with open(tmp_path / "case.bin", "wb") as f:
        for a in (a12, a21, dex, c.b):
            np.ascontiguousarray(a, np.float64).tofile(f)
    print(f"running {DRIVER.name} n={c.n}", flush=True)
    res = subprocess.run([str(DRIVER), str(tmp_path / "faces.bin"), str(tmp_path / "case.bin"),
                          str(tmp_path / "out.bin")], capture_output=True, text=True, timeout=120)
    print(res.stdout[-2000:], res.stderr[-2000:])
    assert res.returncode == 0
    raw = (tmp_path / "out.bin").read_bytes()
    rec = 4 + 16 + 8 * c.n
    assert len(raw) == 11 * rec
    # values assembled on the device, and handed over from a device copy: the
    # same solve as with the host upload (Cg, prec 2 is record 2), bit for bit
    for k in (9, 10):
        assert raw[k * rec:(k + 1) * rec] == raw[2 * rec:3 * rec], k
    tol = 1e-10 * float(np.linalg.norm(c.b))
    k = 0
    for method in ("cg", "bicg", "cgs"):
        for prec in (0, 1, 2):
            blk = raw[k * rec:(k + 1) * rec]
            k += 1
            niter = int(np.frombuffer(blk[:4], np.int32)[0])
            ini, fin = np.frombuffer(blk[4:20], np.float64)
            x = np.frombuffer(blk[20:], np.float64)
            ref = oracle.solve(method, c.row, c.col, c.dia, c.val, c.b, prec=prec, niter=3000,
                               tol=tol)
            slack = max(1, ref["niter"] // 20) + (4 if method == "cgs" else 0)
            assert abs(niter - ref["niter"]) <= slack, (method, prec, niter, ref["niter"])
            assert abs(ini - ref["ini_res"]) <= 1e-12 * ref["ini_res"]
            if ref["niter"] <= 3000:
                assert fin < tol
                assert rel_l2(x, ref["x"]) <= 1e-8, (method, prec)
