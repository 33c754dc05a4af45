"""Trust region of the flexible GMRES (include/stratiflow/NewtonKrylov.h).

With a right preconditioner the update is dx = Z y over the preconditioned
directions Z, which are not orthonormal; the hookstep must bound |dx|, not |y|
(NewtonKrylov.h:196-215 bounds |y| for the orthonormal Krylov basis).
tests/cpp/hookstep_check.cpp takes one Newton step of a linear problem with
preconditioners M^-1 = s I for s = 0 (plain), 1, 10 and 0.1.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_flexible_trust_region_bounds_the_update(tmp_path):
    exe = str(tmp_path / "hookstep_check")
    subprocess.check_call(["g++", "-std=c++14", "-O2", "-DUSE_DOUBLE", "-I" + os.path.join(ROOT, "include"), "-o",
                           exe, os.path.join(ROOT, "tests", "cpp", "hookstep_check.cpp")])
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120, check=True).stdout
    vals = dict((l.split()[0], float(l.split()[1])) for l in out.strip().splitlines())
    Delta = 0.05
    assert vals["full_step_norm"] > 10 * Delta  # the region is active
    for s in ("0", "1", "10", "0.1"):
        n = vals["dx_norm_s" + s]
        assert n <= Delta * (1 + 1e-9), (s, n)
        assert n >= 0.99 * Delta, (s, n)  # and binding
    # inactive region: the flexible solve with M^-1 = 10 I equals the plain one
    assert vals["flex_vs_plain"] < 1e-9
