"""The allocation-free residual (tfs_residual_device_noalloc) through the C ABI:
bitwise the same as tfs_residual_device (and as the oracle's Residual) on two
calls in a row, the second on a smaller matrix; no device allocation (the
library's byte count stays put, where the allocating variant's grows on its
first call); a matrix beyond the scratch capacity is refused untouched."""
import ctypes

import numpy as np
import pytest

from tflows_amd import mesh

pytestmark = pytest.mark.gpu

SCRATCH_ROWS = 1 << 22


@pytest.fixture(scope="module")
def torch():
    import torch as t
    if not t.cuda.is_available():
        pytest.fail("no GPU visible")
    return t


@pytest.fixture(scope="module")
def lib():
    from tflows_amd import solvers
    L = solvers._lib()
    L.tfs_residual_device_noalloc.restype = ctypes.c_int
    L.tfs_residual_device_noalloc.argtypes = [ctypes.c_void_p] * 3
    L.tfs_residual_device.restype = ctypes.c_int
    L.tfs_residual_device.argtypes = [ctypes.c_void_p] * 3
    return L


def _residual(torch, lib, fn, A, x, b):
    r = torch.from_numpy(b.copy()).cuda()
    torch.cuda.synchronize()
    rc = fn(A.handle, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(r.data_ptr()))
    return rc, r.cpu().numpy()


def test_noalloc_residual_matches_the_allocating_one(torch, lib, oracle):
    from tflows_amd import solvers
    cases = [("poisson24_perm", dict(kind=mesh.POISSON, nx=24, beta=1.5, permute=True)),
             ("momentum10", dict(kind=mesh.MOMENTUM, nx=10, beta=2.0, jitter=0.1, permute=True,
                                 pin_first=False, rhs_zero_mean=False))]
    for name, kw in cases:  # the second call is on a smaller matrix
        c = mesh.make_case(**kw)
        print(f"case {name} n={c.n}", flush=True)
        A = solvers.Matrix(c.row, c.col, c.dia, c.val)
        x_h = np.random.default_rng(7).uniform(-1, 1, c.n)
        x = torch.from_numpy(x_h).cuda()
        bytes0 = A.info()["device_bytes"]
        rc, r_na = _residual(torch, lib, lib.tfs_residual_device_noalloc, A, x, c.b)
        assert rc == 0
        assert A.info()["device_bytes"] == bytes0  # nothing allocated
        rc, r_ref = _residual(torch, lib, lib.tfs_residual_device, A, x, c.b)
        assert rc == 0
        assert A.info()["device_bytes"] > bytes0   # the allocating variant's work vectors
        assert np.array_equal(r_na, r_ref)
        assert np.array_equal(r_na, oracle.residual(c.row, c.col, c.dia, c.val, x_h, c.b))
        # and again, now that the matrix's own work vectors exist
        rc, r_na2 = _residual(torch, lib, lib.tfs_residual_device_noalloc, A, x, c.b)
        assert rc == 0 and np.array_equal(r_na2, r_ref)


def test_noalloc_residual_refuses_beyond_capacity(torch, lib):
    from tflows_amd import solvers
    n = SCRATCH_ROWS + 1
    idx = np.arange(n, dtype=np.int32)
    row = np.arange(n + 1, dtype=np.int32)
    A = solvers.Matrix(row, idx, idx, np.full(n, 2.0), base=0)
    x = torch.ones(n, dtype=torch.float64, device="cuda")
    b = np.full(n, 3.0)
    rc, r = _residual(torch, lib, lib.tfs_residual_device_noalloc, A, x, b)
    assert rc == solvers.TFS_ERR_INVALID
    assert np.array_equal(r, b)  # untouched
    rc, r = _residual(torch, lib, lib.tfs_residual_device, A, x, b)
    assert rc == 0 and np.array_equal(r, np.full(n, 1.0))
