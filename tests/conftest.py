import importlib.util
import pathlib
import sys

import numpy as np
import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def load_package():
    """Import t-flows_amd/ under the module name tflows_amd."""
    if "tflows_amd" in sys.modules:
        return sys.modules["tflows_amd"]
    pkg = ROOT / "t-flows_amd"
    spec = importlib.util.spec_from_file_location(
        "tflows_amd", pkg / "__init__.py", submodule_search_locations=[str(pkg)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["tflows_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


load_package()


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")


@pytest.fixture(scope="session")
def oracle():
    from oracle.oracle import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def reference():
    from oracle.oracle import Reference, reference_available
    if not reference_available():
        pytest.skip("reference solver library (oracle/_ref) not built here")
    return Reference()


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def exact_solution(row, col, val, b):
    """x = A^-1 b by sparse LU (scipy), the yardstick for solvers that reach
    the tolerance along a different path than the reference (0-based CRS)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    n = len(row) - 1
    A = sp.csr_matrix((np.asarray(val, np.float64), np.asarray(col), np.asarray(row)), shape=(n, n))
    return spla.spsolve(A.tocsc(), np.asarray(b, np.float64))


def refinement_bar(x_ref_path, x_exact):
    """The accuracy a solve that reaches the tolerance by another path must
    meet: within 10x of the reference path's own error against the exact
    solution, and never worse than rel 1e-8 unless the reference itself is."""
    return max(1e-8, 10.0 * rel_l2(x_ref_path, x_exact))
