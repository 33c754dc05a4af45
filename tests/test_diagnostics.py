"""Diagnostics and field helpers: JoverK (IMEXRK.h:445-459 with the J integrand
of UpdateAdjointVariables, IMEXRK.h:470-490), HorizontalAverage
(Integration.h:37-42), NodalField + Nodal1D.

CPU: the oracle's JoverK against the closed form for single-mode fields,
    B = c + cos(2 pi x / L1) f(z),  W = cos(2 pi x / L1) g(z)
    => J/K = -(1/4) sum_{k=2}^{N3-2} (zf_k - zf_{k-1}) f_k g_k,
independent of the mean c.
GPU: sf_jover_k, sf_field_mean_profile and sf_field_add_profile against the
oracle and numpy on seeded states.
"""
import numpy as np
import pytest

import oracle as O

LX = 2 * np.pi / 0.44


def _single_mode(c, seed, N1=24, N3=32):
    o = O.Oracle(N1, 1, N3, O.TWO_D, LX, 1.0, 10.0, 1000.0, 0.1, 1.0, precision=8)
    rng = np.random.default_rng(seed)
    f, g = rng.standard_normal(N3), rng.standard_normal(N3)
    x = np.arange(N1) / N1
    B = c + np.cos(2 * np.pi * x)[:, None] * f[None, :]
    o.set_nodal("B", B.reshape(1, N1, N3))
    u3 = np.zeros(o.modal_shape, complex)
    u3[0, 1, :] = g / 2  # ToNodal gives cos(2 pi x) g
    o.set_modal("u3", u3)
    z, zf, _, _ = o.grid()
    want = -0.25 * sum((zf[k] - zf[k - 1]) * f[k] * g[k] for k in range(2, N3 - 1))
    return o.jover_k(), want


@pytest.mark.parametrize("c", [0.0, 3.0, -7.5])
def test_oracle_jover_k_closed_form(c):
    got, want = _single_mode(c, 4)
    assert abs(got - want) <= 1e-13 * max(1.0, abs(want)), (got, want)


def test_oracle_jover_k_ignores_mean_buoyancy():
    a, _ = _single_mode(0.0, 5)
    b, _ = _single_mode(10.0, 5)
    assert abs(a - b) < 1e-13 * abs(a)


@pytest.mark.gpu
@pytest.mark.parametrize("dim,N1,N2,N3,prec,tol", [(O.TWO_D, 48, 1, 64, 8, 1e-12), (O.THREE_D, 24, 12, 32, 8, 1e-12),
                                                   (O.TWO_D, 48, 1, 64, 4, 1e-4)])
def test_gpu_jover_k_matches_oracle(dim, N1, N2, N3, prec, tol):
    from helpers import make_pair
    s, o = make_pair(N1, N2, N3, dim, gpu_prec=prec, amplitude=0.05, seed=7)
    s.TimeStep(2)
    o.step(2)
    got, want = s.JoverK(), o.jover_k()
    print("JoverK gpu %.12e oracle %.12e" % (got, want))
    assert abs(got - want) <= tol * abs(want)


@pytest.mark.gpu
def test_gpu_mean_profile_and_add_profile():
    import stratiflow_amd as sf
    from helpers import make_pair
    s, o = make_pair(24, 12, 32, O.THREE_D, gpu_prec=8, amplitude=0.05, seed=9)
    b = o.get_modal("b")
    assert np.allclose(s.state("b").mean_profile(), b[0, 0, :].real, rtol=0, atol=1e-15)
    prof = np.linspace(-1.0, 2.0, 32)
    U = s.field(sf.NODAL, sf.NEUMANN)
    U.upload(o.get_nodal("U1"))
    U.add_profile(prof)
    want = o.get_nodal("U1") + prof[None, None, :]
    assert np.abs(U.download() - want).max() < 1e-14


@pytest.mark.gpu
def test_facade_diagnostics_and_load_and_interpolate(tmp_path):
    """tests/cpp/diag_demo.cpp: LoadAndInterpolate reproduces fields that are
    linear in z exactly at the new points (StateVector.h:275-413); the
    diagnostics of a seeded state match CPU evaluations from the saved file."""
    import os
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    binary = os.path.join(root, "tests", "cpp", "_build", "diag_demo")
    K1, K3, N3 = 48, 64, 96
    L1 = 2 * 3.14159265358979 / 0.44
    old = O.Oracle(K1, 1, K3, O.TWO_D, L1, 1.0, 10.0, 1000.0, 0.1, 1.0, precision=8)
    z, zf, _, _ = old.grid()
    xx = np.arange(K1) / K1
    c, s = np.cos(2 * np.pi * xx)[:, None], np.sin(2 * np.pi * xx)[:, None]
    fields = [c * (0.3 + 0.05 * zf), 0 * c * zf, c * (0.2 - 0.03 * z), 0.1 + 0.02 * zf + s * (0.4 + 0.01 * zf)]
    src = str(tmp_path / "in.fields")
    np.concatenate([f.astype(np.float32).ravel() for f in fields]).tofile(src)
    out, st = str(tmp_path / "out.fields"), str(tmp_path / "state")
    r = subprocess.run([binary, src, out, st], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    vals = dict((l.split()[0], float(l.split()[1])) for l in r.stdout.strip().splitlines())

    new = O.Oracle(K1, 1, N3, O.TWO_D, L1, 1.0, 10.0, 1000.0, 0.1, 1.0, precision=8)
    zn, zfn, _, _ = new.grid()
    got = np.fromfile(out, dtype=np.float32).astype(np.float64).reshape(4, K1, N3)
    want_u1 = c * (0.3 + 0.05 * zfn)
    want_b = 0.1 + 0.02 * zfn + s * (0.4 + 0.01 * zfn)
    assert np.abs(got[0][:, 1:N3 - 1] - want_u1[:, 1:N3 - 1]).max() < 2e-6
    assert np.abs(got[3][:, 1:N3 - 1] - want_b[:, 1:N3 - 1]).max() < 2e-6
    assert np.array_equal(got[0][:, 0], got[0][:, 1]) and np.array_equal(got[0][:, -1], got[0][:, -2])
    want_u3 = c * (0.2 - 0.03 * zn)
    assert np.abs(got[2][:, 3:N3 - 2] - want_u3[:, 3:N3 - 2]).max() < 2e-6
    assert np.abs(got[2][:, [0, 1, 2, N3 - 2, N3 - 1]]).max() == 0.0

    # the seeded state, rebuilt in FP64 from its nodal file
    raw = np.fromfile(st + ".fields", dtype=np.float32).astype(np.float64).reshape(4, 1, K1, N3)
    o = new
    for i, n in ((0, "u1"), (2, "u3"), (3, "b")):
        o.set_modal(n, o.to_modal(raw[i]))
    o.set_nodal("B", raw[3])
    ke, pe = o.energy()
    assert abs(vals["energy"] - (ke + pe)) <= 1e-5 * (ke + pe)
    jk = o.jover_k()
    assert abs(vals["jover_k"] - jk) <= 1e-4 * abs(jk) + 1e-12
    u1, u3, b = o.get_modal("u1"), o.get_modal("u3"), o.get_modal("b")
    A1 = K1 // 3
    w = np.zeros_like(u3)
    for j1 in range(A1):
        col = u1[0, j1]
        dz = o.apply_op("dz_ND", col.real) + 1j * o.apply_op("dz_ND", col.imag)
        w[0, j1] = dz - 1j * (2 * np.pi * j1 / L1) * u3[0, j1]
    ens = o.inner_prod(w, w, 1)
    assert abs(vals["enstrophy"] - ens) <= 1e-5 * ens
    oa = sum((zfn[k] - zfn[k - 1]) * 2 * u3[0, 1, k].imag for k in range(2, N3 - 1))
    assert abs(vals["other_axis"] - oa) <= 1e-5 * abs(oa) + 1e-9
    U = np.tanh(zfn)
    U[1] = U[2]
    U[0] = U[1]
    U[-2] = U[-3]
    U[-1] = U[-2]
    ua, ba = u1[0, 0].real + U, b[0, 0].real
    DY = np.r_[1.0, np.diff(zfn)]
    d = lambda f: np.r_[0.0, 0.0, (f[2:] - f[1:-1]) / DY[2:]]
    with np.errstate(divide="ignore"):
        rig = 0.1 * (d(ba) + 1) / d(ua) ** 2
    assert abs(vals["minimum_ri"] - rig.min()) <= 1e-4 * abs(rig.min())
    ib = L1 * 1.0 * sum((zn[k + 1] - zn[k]) * b[0, 0, k].real for k in range(1, N3 - 1))
    assert abs(vals["integrate_b"] - ib) <= 1e-5 * abs(ib) + 1e-9
    assert abs(vals["background_mid"] - np.tanh(zfn[N3 // 2])) < 1e-5
    assert vals["background_roundtrip"] < 1e-5
    assert os.path.isdir(st + "_plots")


@pytest.mark.gpu
@pytest.mark.parametrize("dim,N1,N2,N3", [(O.TWO_D, 48, 1, 64), (O.THREE_D, 24, 12, 32)])
def test_gpu_field_helpers(dim, N1, N2, N3):
    """PhaseShift (Field.h:1168-1174), RemoveAverage (Integration.h:72-75),
    EnforceBCs per field, CopyToSolver/CopyFromSolver (StateVector.h:511-543)."""
    import stratiflow_amd as sf
    from helpers import make_pair
    s, o = make_pair(N1, N2, N3, dim, gpu_prec=8, amplitude=0.05, seed=11)
    u1 = o.get_modal("u1")
    A1 = N1 // 3
    f = s.field(sf.MODAL, sf.NEUMANN)
    f.upload(u1)
    f.PhaseShift(0.7)
    want = u1.copy()
    want[:, :A1, :] *= np.exp(1j * 0.7 * np.arange(A1))[None, :, None]
    assert np.abs(f.download() - want).max() < 1e-14

    z, zf, _, _ = o.grid()
    for name, bc in (("b", sf.NEUMANN), ("u3", sf.DIRICHLET)):
        a = o.get_modal(name).copy()
        a[0, 0, :] += 0.3  # a mean to remove
        g = s.field(sf.MODAL, bc)
        g.upload(a)
        g.RemoveAverage()
        m = a[0, 0, :].real
        if bc == sf.NEUMANN:
            integral = sum((z[k + 1] - z[k]) * m[k] for k in range(1, N3 - 1))
        else:
            integral = sum((zf[k] - zf[k - 1]) * m[k] for k in range(2, N3 - 1))
        want = a.copy()
        want[0, 0, :] -= integral / 10.0 / 2
        assert np.abs(g.download() - want).max() < 1e-14, name

    # EnforceBCs: Neumann ends copy rows 1 and N3-2, Dirichlet ends are zeroed
    raw = np.random.default_rng(3).standard_normal(o.modal_shape) * (1 + 1j)
    g = s.field(sf.MODAL, sf.NEUMANN)
    g.upload(raw)
    g.EnforceBCs()
    r = g.download()
    act = np.abs(r[..., 5]) > 0
    assert np.array_equal(r[..., 0][act], r[..., 1][act]) and np.array_equal(r[..., -1][act], r[..., -2][act])
    g = s.field(sf.MODAL, sf.DIRICHLET)
    g.upload(raw)
    g.EnforceBCs()
    assert np.abs(g.download()[..., [0, 1, 2, N3 - 2, N3 - 1]]).max() == 0.0

    # CopyToSolver / CopyFromSolver round trip
    names = ["u1", "u2", "u3", "b", "p"]
    src = {}
    for n in names:
        fl = s.field(sf.MODAL, sf.DIRICHLET if n == "u3" else sf.NEUMANN)
        fl.upload(o.get_modal(n) if n != "p" else 0 * u1)
        src[n] = fl
    if dim == O.TWO_D:
        src["u2"] = None
    s.CopyToSolver(*[src[n] for n in names])
    dst = {n: s.field(sf.MODAL, sf.DIRICHLET if n == "u3" else sf.NEUMANN) for n in names}
    s.CopyFromSolver(*[dst[n] for n in names])
    for n in ("u1", "u3", "b"):
        t = s.field(sf.MODAL, sf.DIRICHLET if n == "u3" else sf.NEUMANN)
        t.upload(o.get_modal(n))
        t.EnforceBCs()  # CopyFromSolver ends with EnforceBCs
        assert np.array_equal(dst[n].download(), t.download()), n
