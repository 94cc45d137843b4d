"""bench.py's command-line contract, checked without a GPU: it measures one
GPU and refuses several with a message (and no JSON line); its flags parse."""
import pathlib
import subprocess
import sys

ROOT = pathlib.Path(__file__).resolve().parent.parent


def run(*args):
    return subprocess.run([sys.executable, str(ROOT / "bench.py"), *args], capture_output=True,
                          text=True, timeout=120)


def test_several_gpus_refused():
    r = run("--gpus", "2")
    assert r.returncode != 0
    assert "one GPU" in r.stderr or "single" in r.stderr
    assert r.stdout == ""


def test_help_lists_the_workloads():
    r = run("--help")
    assert r.returncode == 0
    for flag in ("--workload", "--ingest", "--steps", "--warmup", "--inner-order", "--inner-rtol"):
        assert flag in r.stdout


def test_inner_order_names_match_the_library_constants():
    import importlib.util
    from tflows_amd import solvers
    spec = importlib.util.spec_from_file_location("tfs_bench", ROOT / "bench.py")
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    assert bench.INNER_ORDERS == {"reference": solvers.TFS_INNER_REFERENCE,
                                  "coloured": solvers.TFS_INNER_COLOURED,
                                  "bricks": solvers.TFS_INNER_BRICKS,
                                  "auto": solvers.TFS_INNER_AUTO}
