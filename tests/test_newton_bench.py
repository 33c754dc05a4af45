"""The Newton-metric driver (tests/cpp/newton_bench.cpp) on a small sample of
the learned-preconditioner problem class: one training run, one held-out
problem, three Newton steps of each method.

Checks what holds for every method whatever its speed: the residual falls
below the first Newton residual, the learned attempt either meets the GMRES
tolerance or is redone plain (the judge), and the network trained on the GPU
lowers its loss.  bench.py reports the timing comparison at the metric's own
budget.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "_build", "newton_bench_f64")

pytestmark = pytest.mark.gpu


def test_newton_bench_small_sample():
    r = subprocess.run([BIN, "ntrain=1", "ntest=1", "newton=3", "batches=500"], capture_output=True, text=True,
                       timeout=280)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    print(r.stdout)
    tests = {t.split()[4]: t.split() for t in r.stdout.splitlines() if t.startswith("test ")}
    assert set(tests) == {"plain", "learned", "base"}
    for name, t in tests.items():
        steps, rel = int(t[5]), float(t[8])
        assert steps == 3 and rel < 1e-1, (name, t)
    loss = [t.split() for t in r.stdout.splitlines() if t.startswith("loss ")][0]
    assert float(loss[2]) < float(loss[1])
