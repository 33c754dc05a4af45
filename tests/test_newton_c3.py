"""The Newton-Krylov configuration at 256x1x256 (Re=2000, Ri=0.12, Pr=1, tanh U
and background buoyancy, dt=0.01, T=11) in FP32 and FP64 time stepping.

tests/cpp/newton_c3.cpp builds the guess (a seeded perturbation of energy 0.05,
drawn in double in both builds, then 400 plain steps), takes one Newton-GMRES
step with Krylov dimension 60 and trust region Delta = 0.1 (FP64 Krylov algebra
in both), and computes the 10 leading Arnoldi eigenvalues of the linearised
T-map about the updated state (Arnoldi on LinearEvolve, 40 vectors).  The FP64
context is pinned against the FP64 oracle by the parity tests; here the FP32
context is compared with it:
  * the guess and the first Newton residual agree to FP32 rounding growth;
  * the residual after the Newton step agrees to 1e-2 relative;
  * the updated states differ by a few percent: GMRES reaches only ~0.7 of
    the initial residual in 60 vectors, so the hookstep direction follows the
    Krylov space, which the FP32 finite-difference Jacobian perturbs (stated
    bound 0.1);
  * each of the leading FP32 eigenvalues lies within 0.06 max(1, |lambda|) of
    an FP64 one.
The FP64 run is also held against the FP64 oracle's Newton step and
eigenvalues from the same guess (tests/golden/c3_oracle_256x1x256.npz, written
by tests/golden/make_c3_oracle.py): residuals, update norms and the ten
leading eigenvalues.
"""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "_build")
N1, N3 = 256, 256

pytestmark = pytest.mark.gpu


def _run(exe, out):
    os.makedirs(out, exist_ok=True)
    r = subprocess.run([os.path.join(BIN, exe), str(N1), str(N3), "11", "400", "0.05", "61", "1", "1e-8", out, "0.1"],
                       capture_output=True, text=True, timeout=280)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    res, ev = None, []
    for line in r.stdout.splitlines():
        t = line.split()
        if t and t[0] == "newton":
            res = float(t[2])
        elif t and t[0] == "done":
            ratio = float(t[2])
        elif t and t[0] == "ev":
            ev.append(complex(float(t[2]), float(t[3])))
    return res, res * ratio, np.array(ev)


def _fields(path, dtype):
    return np.fromfile(path, dtype=dtype).astype(np.float64).reshape(4, N1, N3)


ORACLE_FIXTURE = os.path.join(ROOT, "tests", "golden", "c3_oracle_256x1x256.npz")


def _against_oracle(out, r0, r1, ev):
    """The FP64 run against the FP64 oracle's Newton step and eigenvalues
    from the same guess (tests/golden/make_c3_oracle.py)."""
    import oracle
    from oracle import krylov
    from stratiflow_amd import ic
    d = np.load(ORACLE_FIXTURE)
    o = oracle.Oracle(N1, 1, N3, oracle.TWO_D, 2 * np.pi / 0.44, 1.0, 10.0, 2000.0, 0.12, 1.0, precision=8)
    z, zf, _, _ = o.grid()
    o.set_background(ic.tanh_background(zf), ic.tanh_stratification(zf))

    def modal(tag):
        raw = np.fromfile(os.path.join(out, tag + ".fields"), dtype=np.float64).reshape(4, 1, N1, N3)
        return krylov.enforce_bcs({"u1": o.to_modal(raw[0]), "u3": o.to_modal(raw[2]), "b": o.to_modal(raw[3])})

    x0, x1 = modal("guess"), modal("solution")
    g = np.array([np.linalg.norm(x0[n]) for n in ("u1", "u3", "b")])
    assert np.allclose(g, d["guess_norms"], rtol=1e-9), "not the guess the fixture was made from"
    dx = krylov.muladd(krylov.copy(x1), -1.0, x0)
    dxn = np.array([np.linalg.norm(dx[n]) for n in ("u1", "u3", "b")])
    cosg = krylov.dot(o, dx, x0, 0.12) / (krylov.norm(o, dx, 0.12) * krylov.norm(o, x0, 0.12))
    print("vs FP64 oracle: residual before %.3e, after %.3e; update norms %s; update.guess %.3e" %
          (abs(r0 - d["r0"]) / d["r0"], abs(r1 - d["r1"]) / d["r1"],
           np.abs(dxn - d["dx_norms"]) / d["dx_norms"], abs(cosg - d["dx_dot_guess"])))
    # the driver prints residuals to 7 digits; the finite-difference Jacobian
    # (step 1e-7 |x|) and 60 Gram-Schmidt steps leave ~1e-7 between the two
    assert abs(r0 - d["r0"]) / d["r0"] < 1e-6
    assert abs(r1 - d["r1"]) / d["r1"] < 1e-5
    assert np.all(np.abs(dxn - d["dx_norms"]) / d["dx_norms"] < 1e-5)
    assert abs(cosg - d["dx_dot_guess"]) < 1e-5
    for k, e in enumerate(d["ev"]):
        dist = np.min(np.abs(ev - e))
        print("eigenvalue %d oracle %.6f%+.6fi  nearest FP64 GPU distance %.3e" % (k, e.real, e.imag, dist))
        assert dist < 1e-5 * max(1.0, abs(e))


def test_newton_step_and_arnoldi_fp32_against_fp64(tmp_path):
    r32, s32, ev32 = _run("newton_c3", str(tmp_path / "f32"))
    r64, s64, ev64 = _run("newton_c3_f64", str(tmp_path / "f64"))
    print("residual before: fp32 %.9e fp64 %.9e; after one step: fp32 %.9e fp64 %.9e" % (r32, r64, s32, s64))
    assert abs(r32 - r64) / r64 < 1e-3
    assert abs(s32 - s64) / s64 < 1e-2
    assert s64 < r64
    for tag, bound in (("guess", 1e-3), ("solution", 0.1)):
        a = _fields(str(tmp_path / "f32" / (tag + ".fields")), np.float32)
        b = _fields(str(tmp_path / "f64" / (tag + ".fields")), np.float64)
        for i, n in ((0, "U1"), (2, "U3"), (3, "B")):
            e = np.linalg.norm(a[i] - b[i]) / np.linalg.norm(b[i])
            print("%s %s fp32 vs fp64 rel l2 %.3e" % (tag, n, e))
            assert e < bound, (tag, n, e)
    assert len(ev32) == 10 and len(ev64) == 10
    if os.path.exists(ORACLE_FIXTURE):
        _against_oracle(str(tmp_path / "f64"), r64, s64, ev64)
    for k, e in enumerate(ev32):
        d = np.min(np.abs(ev64 - e))
        print("eigenvalue %d fp32 %.6f%+.6fi  nearest fp64 distance %.3e" % (k, e.real, e.imag, d))
        assert d < 0.06 * max(1.0, abs(e))
    assert abs(abs(ev32[0]) - abs(ev64[0])) / abs(ev64[0]) < 0.03
