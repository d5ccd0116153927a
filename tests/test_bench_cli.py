"""bench.py's command line on the host: one GPU only (a message and a non-zero exit for
--gpus > 1, before any device work), and the byte counts behind its roofline figures."""
import importlib.util
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2", "--steps", "1"],
                       capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert r.returncode != 0
    assert "one GPU" in r.stderr
    assert r.stdout.strip() == ""


def test_bytes_per_node_stage():
    """LSRK54 over its five stages: FP64 reads q and writes du and q in stage 1 (96 B; its du term
    has a_1 = 0), and also reads du in stages 2-5 (128 B); mixed keeps du in FP32."""
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    b = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(b)
    assert abs(b.BYTES_FP64 - 121.6) < 1e-12
    assert abs(b.BYTES_MIXED - 92.8) < 1e-12
    assert abs(b.BYTES_FP64 - (4 * 8 * 3 + 4 * (4 * 8 * 4)) / 5) < 1e-12
    assert abs(b.BYTES_MIXED - ((4 * 8 * 2 + 4 * 4) + 4 * (4 * 8 * 2 + 4 * 4 * 2)) / 5) < 1e-12
    assert b.HBM_PEAK_GBS == 8000.0
