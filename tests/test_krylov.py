"""Newton-GMRES and Arnoldi drivers (NewtonKrylov.h, Arnoldi.h).

CPU: the FP64 dense helpers (include/stratiflow/dense.h: Jacobi SVD, hookstep
least squares, Hessenberg QR eigenvalues) against numpy's LAPACK routines, and
the oracle's Python GMRES/Arnoldi restatement against closed forms on small
dense operators.
GPU: the C++ drivers through the facade (FP64 context) against the oracle
restatement on the same seeded 2D problem.
"""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
KRYLOV_BIN = os.path.join(ROOT, "tests", "cpp", "_build", "krylov_demo")


@pytest.fixture(scope="module")
def dense_out(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("dense") / "dense_check")
    r = subprocess.run(["g++", "-std=c++14", "-O2", "-Wall", "-Wextra", "-Werror", "-I" + os.path.join(ROOT, "include"),
                        "-o", exe, os.path.join(ROOT, "tests", "cpp", "dense_check.cpp")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout
    cases = {}
    for line in out.splitlines():
        t = line.split()
        key, c = t[0], int(t[1])
        d = cases.setdefault(c, {"LS": [], "EVEC": []})
        if key == "M":
            r_, c_ = int(t[2]), int(t[3])
            d["M"] = np.array([float(v) for v in t[4:]]).reshape(c_, r_).T
        elif key == "SV":
            d["SV"] = np.array([float(v) for v in t[2:]])
        elif key in ("REC", "ORTH"):
            d[key] = float(t[2])
        elif key == "LS":
            d["LS"].append((float(t[2]), float(t[3]), float(t[4]), np.array([float(v) for v in t[5:]])))
        elif key == "EV":
            v = np.array([float(x) for x in t[2:]])
            d["EV"] = v[0::2] + 1j * v[1::2]
        elif key == "EVEC":
            d["EVEC"].append(float(t[3]))
    return cases


def test_jacobi_svd_matches_lapack(dense_out):
    n = 0
    for c, d in dense_out.items():
        if "SV" not in d:
            continue
        ref = np.linalg.svd(d["M"], compute_uv=False)
        assert np.allclose(d["SV"], ref, rtol=1e-12, atol=1e-13), c
        assert d["REC"] < 1e-12 and d["ORTH"] < 1e-12
        n += 1
    assert n == 5


def test_hookstep_least_squares(dense_out):
    """Delta large: the plain least-squares solution; Delta small: |y| = |z| <= Delta, mu > 0."""
    from oracle import krylov
    for c, d in dense_out.items():
        if not d["LS"]:
            continue
        H = d["M"]
        beta = 2.0
        for Delta, res, mu, y in d["LS"]:
            rhs = np.zeros(H.shape[0])
            rhs[0] = beta
            yr, resr, mur = krylov.hookstep(H, beta, Delta)
            assert abs(mu - mur) < 1e-9
            assert np.allclose(y, yr, rtol=1e-9, atol=1e-11), (c, Delta)
            assert abs(res - resr) < 1e-10
            if mu == 0:
                ls = np.linalg.lstsq(H, rhs, rcond=None)[0]
                assert np.allclose(y, ls, rtol=1e-9, atol=1e-11)
            else:
                assert np.linalg.norm(y) <= Delta * (1 + 1e-12)
                # mu is the first point of the 1e-5 grid whose hookstep fits
                # (NewtonKrylov.h:196-203): z(mu) = p d/(d^2 + mu) in the singular basis
                U, sv, _ = np.linalg.svd(H, full_matrices=False)
                p = U[0, :] * beta
                z = lambda m: p * sv / (sv * sv + m)
                assert np.linalg.norm(z(mur)) <= Delta * (1 + 1e-12)
                assert np.linalg.norm(z(mur - 1e-5)) > Delta
                assert abs(mur / 1e-5 - round(mur / 1e-5)) < 1e-3  # sums of 1e-5 in double


def test_hessenberg_eigenvalues_match_lapack(dense_out):
    n = 0
    for c, d in dense_out.items():
        if "EV" not in d:
            continue
        ref = np.linalg.eigvals(d["M"])
        got = d["EV"]
        assert len(got) == len(ref)
        # match each reference eigenvalue to the nearest computed one
        for e in ref:
            assert np.min(np.abs(got - e)) < 1e-9 * max(1.0, abs(e)), (c, e)
        for r in d["EVEC"]:
            assert r < 1e-8
        n += 1
    assert n == 6


def _dense_problem(n=12, seed=3):
    rng = np.random.default_rng(seed)
    A = np.eye(n) + 0.3 * rng.standard_normal((n, n))
    b = rng.standard_normal(n)
    return A, b


class _DenseOracle:
    """A stand-in for the oracle's inner product on plain vectors (all weights 1)."""
    modal_shape = (1, 1, 1)

    def inner_prod(self, a, b, bc):
        return float(np.real(np.vdot(b, a)))


def _vec(x):
    # one 'u1' field holding a column with 2 padding slots on each end, so that
    # EnforceBCs (NeumannEnds) leaves the interior untouched
    a = np.zeros((1, 1, len(x) + 4), complex)
    a[0, 0, 2:-2] = x
    return {"u1": a}


def test_oracle_gmres_solves_dense_system():
    """The restated GMRES (finite differences of an affine F) solves A dx = F(x) to its tolerance."""
    from oracle import krylov
    A, b = _dense_problem()
    n = len(b)
    o = _DenseOracle()

    def F(v):  # F(x) = b - A x, so -dF/dx = A and the Newton update solves A dx = F(x)
        x = v["u1"][0, 0, 2:-2].real
        return krylov.enforce_bcs(_vec(b - A @ x))

    x0 = _vec(np.zeros(n))
    x0["u1"][0, 0, 2:-2] = 0.1
    x0 = krylov.enforce_bcs(x0)
    Fx = F(x0)
    dx, H, hist = krylov.gmres(o, F, x0, Fx, 0.0, n + 1, 1e-10, 1e6, fd_epsilon=1e-6)
    xs = x0["u1"][0, 0, 2:-2].real + dx["u1"][0, 0, 2:-2].real
    assert np.allclose(A @ xs, b, atol=1e-6)
    # GMRES residuals never increase
    assert all(hist[i + 1] <= hist[i] * (1 + 1e-9) for i in range(len(hist) - 1))


def test_oracle_arnoldi_eigenvalues_of_dense_map():
    """Full-dimension Arnoldi of a dense map recovers all its eigenvalues."""
    from oracle import krylov
    rng = np.random.default_rng(5)
    n = 10
    A = rng.standard_normal((n, n))
    o = _DenseOracle()

    def L(v):
        x = v["u1"][0, 0, 2:-2].real
        return krylov.enforce_bcs(_vec(A @ x))

    q0 = krylov.enforce_bcs(_vec(rng.standard_normal(n)))
    ev, H, q = krylov.arnoldi(o, L, q0, 0.0, n + 1)
    ref = np.linalg.eigvals(A)
    for e in ref:
        assert np.min(np.abs(ev - e)) < 1e-8 * max(1, abs(e))


# ---------------------------------------------------------------------------
# GPU: the C++ drivers against the restatement
# ---------------------------------------------------------------------------
def _read_fields(path, o):
    N1, N2, N3 = o.N1, o.N2, o.N3
    raw = np.fromfile(path, dtype=np.float64).reshape(4, N2, N1, N3)
    return {"u1": o.to_modal(raw[0]), "u3": o.to_modal(raw[2]), "b": o.to_modal(raw[3])}


@pytest.mark.gpu
def test_newton_gmres_and_arnoldi_match_oracle(tmp_path):
    import oracle
    from oracle import krylov
    from stratiflow_amd import ic
    K = 9
    r = subprocess.run([KRYLOV_BIN, str(tmp_path), str(K)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    res, gres, H, ev = {}, {}, np.zeros((K, K - 1)), []
    for line in r.stdout.splitlines():
        t = line.split()
        if t[0] == "res":
            res[int(t[1])] = float(t[2])
        elif t[0] == "gres":
            gres[int(t[1])] = float(t[2])
        elif t[0] == "H":
            H[int(t[1]), int(t[2])] = float(t[3])
        elif t[0] == "ev":
            ev.append(float(t[2]) + 1j * float(t[3]))
    ev = np.array(ev)

    N1, N2, N3 = 32, 1, 48
    L1 = 2 * np.pi / 0.44
    Ri = 0.1
    o = oracle.Oracle(N1, N2, N3, oracle.TWO_D, L1, 1.0, 10.0, 100.0, Ri, 1.0, precision=8)
    z, zf, _, _ = o.grid()
    o.set_background(ic.tanh_background(zf))
    x0 = krylov.enforce_bcs(_read_fields(str(tmp_path / "x0.fields"), o))

    def F(v):
        out = krylov.phi_fixed(o, v, 0.02, 25)
        return krylov.muladd(out, -1.0, v)

    Fx = F(x0)
    r0 = krylov.norm(o, Fx, Ri)
    print("Newton residual gpu %.12e oracle %.12e" % (res[0], r0))
    assert abs(res[0] - r0) / r0 < 1e-9
    dx, Ho, hist = krylov.gmres(o, F, x0, Fx, Ri, K, 1e-14, 100.0, fd_epsilon=1e-7)
    eH = np.abs(H - Ho).max() / np.abs(Ho).max()
    print("Hessenberg rel diff %.3e" % eH)
    assert eH < 1e-5
    for k, g in gres.items():
        assert abs(g - hist[k]) <= 1e-5 * max(hist[k], 1e-12) + 1e-9
    x1o = krylov.muladd(krylov.copy(x0), 1.0, dx)
    x1 = _read_fields(str(tmp_path / "x1.fields"), o)
    e1 = max(np.linalg.norm(x1[n] - x1o[n]) / np.linalg.norm(x1o[n]) for n in x1)
    print("Newton update rel diff %.3e" % e1)
    assert e1 < 1e-5
    # the Newton step reduced the residual
    assert res[1] < res[0]

    # Arnoldi of LinearEvolve (T = 0.5, dt = 0.01) about x0
    q0 = krylov.enforce_bcs(_read_fields(str(tmp_path / "q0.fields"), o))
    evo, Ha, _ = krylov.arnoldi(o, lambda v: krylov.linear_evolve(o, v, x0, 0.5), q0, Ri, K + 4)
    print("eigenvalues gpu", ev[:4], "oracle", evo[:4])
    for e in evo[:6]:
        assert np.min(np.abs(ev - e)) < 1e-6 * max(1.0, abs(e))
