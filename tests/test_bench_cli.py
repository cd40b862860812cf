"""bench.py without a GPU: every function main() calls is defined, the
multi-GPU request exits with a message, and the committed counter summaries
are read back (traffic ratio per kernel)."""
import ast
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench_source():
    return open(os.path.join(ROOT, "bench.py")).read()


def test_every_called_module_function_is_defined():
    tree = ast.parse(_bench_source())
    defined = {n.name for n in tree.body if isinstance(n, ast.FunctionDef)}
    main = next(n for n in tree.body if isinstance(n, ast.FunctionDef) and n.name == "main")
    called = {n.func.id for n in ast.walk(main) if isinstance(n, ast.Call) and isinstance(n.func, ast.Name)}
    local = {n.id for n in ast.walk(main) if isinstance(n, ast.Name) and isinstance(n.ctx, ast.Store)}
    local |= {n.name for n in ast.walk(main) if isinstance(n, ast.FunctionDef)}
    local |= {a.asname or a.name for n in ast.walk(main) if isinstance(n, ast.ImportFrom) for a in n.names}
    import builtins as _b
    builtins = set(dir(_b))
    missing = {c for c in called if c not in defined and c not in local and c not in builtins}
    assert not missing, missing
    for name in ("batched_line", "cpu_baseline", "traffic_probe", "roofline_entry"):
        assert name in defined


def test_more_than_one_gpu_exits_with_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)


def test_traffic_ratio_read_from_profiles():
    sys.path.insert(0, ROOT)
    import bench
    assert abs(bench.measured_traffic_ratio("sbr_chase") - 0.3309) < 1e-12
    assert abs(bench.measured_traffic_ratio("sbr_symm") - 2.1711) < 1e-12
    assert abs(bench.measured_traffic_ratio("pcg_symv") - 1.1263) < 1e-12
    assert abs(bench.measured_traffic_ratio("scs_build") - 1.0197) < 1e-12
    assert bench.measured_traffic_ratio("trd_symv") is None
    assert bench.measured_traffic_ratio("no_such_kernel") is None
