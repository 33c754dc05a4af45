"""Diffusion preconditioner (sf_field_diffusion_precondition) against what the
mathematics fixes: on modes with |k| <= kmax, y = out + in/2 must satisfy
  tnu*(Dzz - k^2) y - tadv*U(z)*ddx(y) - delta*y = in
on the interior rows (ddx(y) = i kx y: the advection term -i tadv kx U y), with Dzz formed independently by two applications of
the library's first-derivative operators (Dz(D->N) o Dz(N->D) for Neumann
fields, Dz(N->D) o Dz(D->N) for Dirichlet: the same stencils as the
Neumannified/Dirichletified second derivative, Differentiation.cpp:7-160) and
-k^2 by two applications of ddx (and ddy); the boundary rows satisfy the
Neumannify/Dirichletify rows; every other mode gives out = -in exactly."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

LX = 2 * np.pi / 0.44


def _random_modal(s, seed):
    rng = np.random.default_rng(seed)
    M1 = s.N1 // 2 + 1
    return rng.standard_normal((s.N2, M1, s.N3)) + 1j * rng.standard_normal((s.N2, M1, s.N3))


@pytest.mark.parametrize("prec,tol", [(8, 1e-11), (4, 2e-4)])
@pytest.mark.parametrize("shape", [(48, 1, 64, 1), (24, 12, 100, 0)])
@pytest.mark.parametrize("bc", [0, 1])
@pytest.mark.parametrize("tadv", [0.0, 2.0])
def test_solves_the_diffusion_system(shape, bc, prec, tol, tadv):
    import stratiflow_amd as sf
    from stratiflow_amd import ic
    L = sf.lib()
    N1, N2, N3, dim = shape
    print("precond", shape, "bc", bc, "prec", prec, "tadv", tadv, flush=True)
    s = sf.IMEXRK(N1, N2, N3, dim, LX, LX / 2, 10.0, 1000.0, 0.1, 1.0, precision=prec)
    _, zf = s.vertical_points()
    U = ic.tanh_background(zf)
    s.SetBackground(U)
    other = sf.DIRICHLET if bc == sf.NEUMANN else sf.NEUMANN
    v = s.field(sf.MODAL, bc)
    v.upload(_random_modal(s, 5))
    vin = v.download()
    out = s.field(sf.MODAL, bc)
    tnu, delta, kmax = 0.02, 1e-2, 1.0
    v.diffusion_precondition(out, tnu, delta, kmax, tadv)
    o = out.download()
    y_f = s.field(sf.MODAL, bc)
    y_f.upload(o + 0.5 * vin)
    y = y_f.download()
    d1 = s.field(sf.MODAL, other)
    d2 = s.field(sf.MODAL, bc)
    s._chk(L.sf_ddz(y_f.h, d1.h))
    s._chk(L.sf_ddz(d1.h, d2.h))
    dzz = d2.download()
    kk = s.field(sf.MODAL, bc)
    t1 = s.field(sf.MODAL, bc)
    s._chk(L.sf_ddx(y_f.h, t1.h))
    dxy = t1.download()
    s._chk(L.sf_ddx(t1.h, kk.h))
    k2y = kk.download()
    if dim == 0:
        s._chk(L.sf_ddy(y_f.h, t1.h))
        s._chk(L.sf_ddy(t1.h, kk.h))
        k2y = k2y + kk.download()
    kx = 0.44 * np.arange(N1 // 2 + 1)
    ky = (2 * np.pi / (LX / 2)) * np.fft.fftfreq(N2, 1.0 / N2) if dim == 0 else np.zeros(1)
    k2 = ky[:, None] ** 2 + kx[None, :] ** 2
    inner = k2 <= kmax ** 2
    active = np.abs(vin).sum(axis=2) > 0
    assert active[inner].any() and active[~inner].any()
    lo = 3 if bc == sf.DIRICHLET else 2
    sl = slice(lo, N3 - 2)
    lhs = tnu * (dzz + k2y) - tadv * U * dxy - delta * y
    scale = np.abs(tnu * dzz[..., sl]).max() + np.abs(vin[..., sl]).max()
    sel = inner & active
    err = np.abs(lhs[sel][:, sl] - vin[sel][:, sl]).max() / scale
    print("  interior residual %.3e" % err, flush=True)
    assert err < tol
    ys = y[sel]
    ysc = np.abs(ys).max()
    if bc == sf.NEUMANN:
        ends = [ys[:, 0], ys[:, 1] - ys[:, 2], ys[:, N3 - 3] - ys[:, N3 - 2], ys[:, N3 - 1]]
    else:
        ends = [ys[:, 0], ys[:, 1], ys[:, 2], ys[:, N3 - 2], ys[:, N3 - 1]]
    eend = max(np.abs(e).max() for e in ends) / ysc
    print("  boundary rows %.3e" % eend, flush=True)
    assert eend < tol
    outer = (~inner) & active
    assert np.array_equal(o[outer], -vin[outer])
    s.close()


def test_rejects_bad_arguments():
    import stratiflow_amd as sf
    s = sf.IMEXRK(48, 1, 64, sf.TWO_D, LX, 1.0, 10.0, 1000.0, 0.1, 1.0)
    a = s.field(sf.MODAL, sf.NEUMANN)
    b = s.field(sf.MODAL, sf.DIRICHLET)
    with pytest.raises(sf.SFError):
        a.diffusion_precondition(b, 0.02, 1e-2)  # BCs differ
    with pytest.raises(sf.SFError):
        a.diffusion_precondition(a, 0.02, 1e-2)  # in place
    c = s.field(sf.MODAL, sf.NEUMANN)
    with pytest.raises(sf.SFError):
        a.diffusion_precondition(c, 0.02, 0.0)  # delta must be positive
    s.close()


@pytest.mark.gpu
def test_advected_preconditioner_beats_plain_gmres():
    """Newton-GMRES on the held-out problems of the learned-preconditioner demo
    (64x1x64, Re = 100, T = 2, Krylov dimension 61, tolerance 2e-2): with the
    advected diffusion preconditioner (delta = 0.3, every mode) GMRES needs at
    most 0.75x the operator applications of plain GMRES in total and converges
    on at least as many problems (tests/cpp/precond_demo.cpp)."""
    import os
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    binary = os.path.join(root, "tests", "cpp", "_build", "precond_demo")
    r = subprocess.run([binary, "4", "61", "2e-2", "100", "0.3", "100", "1"], capture_output=True, text=True,
                       timeout=200)
    assert r.returncode == 0, r.stdout + r.stderr
    print(r.stdout, flush=True)
    total = [line.split() for line in r.stdout.splitlines() if line.startswith("total")]
    assert len(total) == 1
    plain_mv, pre_mv, plain_ok, pre_ok = (int(x) for x in total[0][1:5])
    assert pre_mv <= 0.75 * plain_mv, (pre_mv, plain_mv)
    assert pre_ok >= plain_ok, (pre_ok, plain_ok)
