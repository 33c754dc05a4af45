"""Pins for the CPU oracle (oracle/sf_oracle.cpp) that do not use the oracle's own formulas.

The reference ships no tests, fixtures or golden vectors (it has no test
directory; Catch is vendored but unused), so every function of the oracle is
checked against mathematics instead:
  * grid            -- closed-form end conditions and symmetry (Field.cpp:3-31)
  * operators       -- exactness on linear data, adjoint identities under the
                       reference's quadrature weights, composition identities
  * FFT             -- numpy.fft (pocketfft), an independent implementation
  * Thomas          -- numpy.linalg.solve on the dense matrix
  * Crank-Nicolson  -- closed-form amplification of a vertically uniform mode
  * time stepping   -- second-order convergence under dt halving
  * projection      -- discrete divergence vanishes on interior rows
  * InnerProd       -- Parseval in 2D; the documented j1=0 double count in 3D
  * physics         -- Kelvin-Helmholtz growth rate of the tanh shear layer
"""
import numpy as np
import pytest

import oracle as O

LX = 2 * np.pi / 0.44


def make(N1=24, N2=1, N3=32, dim=O.TWO_D, Re=1000.0, Ri=0.1, Pr=1.0, prec=8, L3=10.0):
    return O.Oracle(N1, N2, N3, dim, LX, LX / 2, L3, Re, Ri, Pr, precision=prec)


# ---------------------------------------------------------------- grid
@pytest.mark.parametrize("N", [16, 64, 384])
def test_grid_closed_form(N):
    o = make(N3=N)
    z, zf, DY, DYF = o.grid()
    L = 10.0
    # walls sit on the Neumann points 1 and N-2 (Field.cpp:14-16)
    assert zf[1] == pytest.approx(-L, abs=1e-12)
    assert zf[N - 2] == pytest.approx(L, abs=1e-12)
    # antisymmetric stretched grid
    assert np.allclose(z[1:], -z[1:][::-1], atol=1e-12)
    assert np.all(np.diff(z[1:]) > 0) and np.all(DY[1:] > 0) and np.all(DYF > 0)
    # refinement at the centre: x^7 map is flattest at x = 0
    assert DYF[N // 2] < DYF[2]


@pytest.mark.parametrize("N", [16, 64, 384])
def test_grid_is_the_degree7_map(N):
    """z_1..z_{N-1} over x = linspace(-1, 1, N-1) is an odd degree-7 polynomial
    s*(2x^7 + x)/3 (Field.cpp:9-13): a least-squares fit recovers the x^7 : x
    coefficient ratio 2 and nothing else.  The older cubic map of Field.cpp:33-48
    (2x^3 + x)/3 and a dropped or changed power would put weight on other
    coefficients or change the ratio."""
    z, zf, DY, DYF = make(N3=N).grid()
    x = np.linspace(-1.0, 1.0, N - 1)
    c = np.polynomial.polynomial.polyfit(x, z[1:], 7)
    s = 3 * c[1]
    assert abs(c[7] / c[1] - 2.0) < 1e-9
    for k in (0, 2, 3, 4, 5, 6):
        assert abs(c[k]) < 1e-9 * abs(s)
    # the scale puts the first Neumann point (z_1 + z_2)/2 on the wall -L
    # (Field.cpp:15-17)
    assert 0.5 * (z[1] + z[2]) == pytest.approx(-10.0, abs=1e-12)


def test_grid_spacing_ratio_centre_to_wall():
    """dz/dx of s*(2x^7 + x)/3 is 1 at the centre and 15 at the walls, so the
    centre-to-wall spacing ratio tends to 1/15 as N grows (the cubic map of
    Field.cpp:33-48 gives 1/7)."""
    ratios = []
    for N in (128, 512, 2048):
        z, zf, DY, DYF = make(N3=N).grid()
        ratios.append(DYF[N // 2] / DYF[1])
    assert abs(ratios[-1] - 1.0 / 15.0) < 2e-3
    assert abs(ratios[-1] - 1.0 / 15.0) < abs(ratios[0] - 1.0 / 15.0)


# ---------------------------------------------------------------- CFL
def _centre_gap(N, L=10.0):
    """z(N/2+1) - z(N/2) of the x^7 map in closed form: the points sit at
    x = 0 and x = 2/(N-2) of linspace(-1, 1, N-1), scaled so that the first
    Neumann point lies on -L (Field.cpp:3-19)."""
    h = lambda x: L * (2 * x ** 7 + x) / 3
    x1, x2 = -1.0, -1.0 + 2.0 / (N - 2)
    scale = -L / (0.5 * (h(x1) + h(x2)))
    return scale * (h(2.0 / (N - 2)) - h(0.0))


@pytest.mark.parametrize("dim,N2", [(O.TWO_D, 1), (O.THREE_D, 12)])
def test_cfl_uniform_flow_closed_form(dim, N2):
    """IMEXRK::CFL (IMEXRK.h:355-374) for uniform nodal velocities U1 = c,
    U2 = v, U3 = w and a uniform background U_ = a:
      dt * (|c + a| N1/L1 + |v| N2/L2 + |w| / (z(N3/2+1) - z(N3/2)))."""
    N1, N3, dt = 24, 48, 0.013
    o = make(N1=N1, N2=N2, N3=N3, dim=dim)
    o.set_dt(dt)
    c, a, v, w = 0.5, 0.25, -0.7, -0.3
    o.set_background(np.full(N3, a))
    shape = o.nodal_shape
    o.set_nodal("U1", np.full(shape, c))
    o.set_nodal("U3", np.full(shape, w))
    if dim == O.THREE_D:
        o.set_nodal("U2", np.full(shape, v))
    d3 = _centre_gap(N3)
    want = dt * (abs(c + a) * N1 / LX + (abs(v) * N2 / (LX / 2) if dim == O.THREE_D else 0.0) + abs(w) / d3)
    assert o.cfl() == pytest.approx(want, rel=1e-12)
    # the largest |value| wins, wherever it sits (here a single negative spike)
    U3 = np.full(shape, w)
    U3[0, N1 // 3, 5] = -2.0
    o.set_nodal("U3", U3)
    want2 = want + dt * (2.0 - abs(w)) / d3
    assert o.cfl() == pytest.approx(want2, rel=1e-12)


@pytest.mark.parametrize("dim,N2", [(O.TWO_D, 1), (O.THREE_D, 12)])
def test_cfl_linear_uniform_closed_form(dim, N2):
    """IMEXRK::CFLlinear (IMEXRK.h:376-393) about a uniform linearisation state
    u1_tot = c, u2_tot = v, u3_tot = w (the (0,0) mode, so every nodal point
    holds the value):  dt * ((1 + |c|) N1/L1 + |v| N2/L2 + |w|/dz_mid); the
    background U_ does not enter (1 bounds it)."""
    N1, N3, dt = 24, 48, 0.02
    o = make(N1=N1, N2=N2, N3=N3, dim=dim)
    o.set_dt(dt)
    o.set_background(np.full(N3, 5.0))  # must not matter
    c, v, w = -0.4, 0.6, 0.35

    def uniform(val):
        m = np.zeros(o.modal_shape, dtype=complex)
        m[0, 0, :] = val
        return m
    o.set_modal("u1t", uniform(c))
    o.set_modal("u3t", uniform(w))
    o.set_modal("bt", uniform(0.0))
    if dim == O.THREE_D:
        o.set_modal("u2t", uniform(v))
    o.set_linear_background()
    d3 = _centre_gap(N3)
    want = dt * ((1 + abs(c)) * N1 / LX + (abs(v) * N2 / (LX / 2) if dim == O.THREE_D else 0.0) + abs(w) / d3)
    assert o.cfl(linear=True) == pytest.approx(want, rel=1e-12)


# ---------------------------------------------------------------- operators
def test_operators_exact_on_linear_data():
    o = make(N3=48)
    z, zf, DY, DYF = o.grid()
    a, b = 0.3, -1.7
    fN, fD = a + b * zf, a + b * z
    N = len(z)
    # Neumann -> Dirichlet derivative rows 2..N-1
    assert np.allclose(o.apply_op("dz_ND", fN)[2:], b)
    assert np.all(o.apply_op("dz_ND", fN)[:2] == 0)
    # Dirichlet -> Neumann derivative rows 1..N-2
    assert np.allclose(o.apply_op("dz_DN", fD)[1:N - 1], b)
    # full interpolation onto Dirichlet points is exact for linear data on
    # rows 2..N-2; rows 1 and N-1 use the copied ghost spacings DYF_0 = DYF_1
    # and DYF_{N-1} = DYF_{N-2} of dzFractional (Field.cpp:79-80)
    assert np.allclose(o.apply_op("ifull", fN)[2:N - 1], fD[2:N - 1])
    assert not np.isclose(o.apply_op("ifull", fN)[1], fD[1])
    # Dirichlet -> Neumann midpoint interpolation is exact for linear data
    # (row 0 would read the uninitialised VerticalPoints(0), Field.cpp:11)
    assert np.allclose(o.apply_op("idir", fD)[1:N - 1], fN[1:N - 1])
    # both second derivatives annihilate linear data
    assert np.allclose(o.apply_op("dzz_N", fN)[1:N - 1], 0, atol=1e-9)
    assert np.allclose(o.apply_op("dzz_D", fD)[2:N - 1], 0, atol=1e-9)
    # tilde and bar interpolations preserve constants
    assert np.allclose(o.apply_op("itilde", np.ones(N))[2:N - 1], 1)
    assert np.allclose(o.apply_op("ibar", np.ones(N))[1:], 1)


def test_operator_adjoints_and_compositions():
    o = make(N3=40)
    z, zf, DY, DYF = o.grid()
    N = len(z)
    rng = np.random.default_rng(3)
    f = np.zeros(N)
    g = np.zeros(N)
    f[3:N - 3] = rng.standard_normal(N - 6)
    g[3:N - 3] = rng.standard_normal(N - 6)
    # Itilde is the adjoint of Idir under the weights DY (Dirichlet) and DYF (Neumann)
    lhs = np.sum(DY * g * o.apply_op("itilde", f))
    rhs = np.sum(DYF * f * o.apply_op("idir", g))
    assert lhs == pytest.approx(rhs, rel=1e-12)
    # Dz(N->D) = -Dz(D->N)^* under the same weights
    lhs = np.sum(DY * g * o.apply_op("dz_ND", f))
    rhs = -np.sum(DYF * f * o.apply_op("dz_DN", g))
    assert lhs == pytest.approx(rhs, rel=1e-12)
    # second derivatives are compositions of the first derivatives
    full = rng.standard_normal(N)
    dzzN = o.apply_op("dzz_N", full)
    comp = o.apply_op("dz_DN", o.apply_op("dz_ND", full))
    assert np.allclose(dzzN[2:N - 2], comp[2:N - 2], rtol=1e-10, atol=1e-8)
    dzzD = o.apply_op("dzz_D", full)
    comp = o.apply_op("dz_ND", o.apply_op("dz_DN", full))
    assert np.allclose(dzzD[2:N - 2], comp[2:N - 2], rtol=1e-10, atol=1e-8)


# ---------------------------------------------------------------- FFT
def _mask(N1, N2, three):
    M1 = N1 // 2 + 1
    m = np.zeros((N2, M1), bool)
    for j2 in range(N2):
        for j1 in range(M1):
            ok = j1 < N1 // 3
            if three and N2 > 1:
                ok = ok and (j2 < N2 // 3 or j2 >= N2 - N2 // 3 + 1)
            m[j2, j1] = ok
    return m


@pytest.mark.parametrize("N1,N2,dim", [(12, 1, O.TWO_D), (48, 1, O.TWO_D), (12, 6, O.THREE_D), (24, 18, O.THREE_D)])
@pytest.mark.parametrize("prec,tol", [(8, 1e-13), (4, 2e-6)])
def test_fft_matches_numpy(N1, N2, dim, prec, tol):
    N3 = 16
    o = make(N1=N1, N2=N2, N3=N3, dim=dim, prec=prec)
    rng = np.random.default_rng(5)
    x = rng.standard_normal((N2, N1, N3))
    mask = _mask(N1, N2, dim == O.THREE_D)
    ref = np.fft.rfft2(x, axes=(0, 1)) / (N1 * N2) * mask[:, :, None]
    got = o.to_modal(x)
    assert np.abs(got - ref).max() < tol * np.abs(ref).max()
    back = o.to_nodal(got)
    refb = np.fft.irfft2(ref, s=(N2, N1), axes=(0, 1)) * (N1 * N2)
    assert np.abs(back - refb).max() < tol * np.abs(refb).max() * 10


def test_c2r_ignores_imaginary_dc():
    o = make(N1=12, N2=1, N3=8)
    m = np.zeros((1, 7, 8), complex)
    m[0, 0, :] = 1.0 + 5.0j
    nod = o.to_nodal(m)
    assert np.allclose(nod, 1.0)


# ---------------------------------------------------------------- Thomas
@pytest.mark.parametrize("kind", [0, 1, 2, 3])
@pytest.mark.parametrize("j1", [0, 3])
def test_thomas_matches_dense_solve(kind, j1):
    N3 = 48
    o = make(N1=24, N3=N3, Re=300.0, Pr=0.7)
    o.set_dt(0.05)
    rng = np.random.default_rng(kind * 10 + j1)
    d = rng.standard_normal(N3) + 1j * rng.standard_normal(N3)
    x, (a, b, c) = o.tri_solve(kind, 1, j1, 0, d)
    A = np.diag(b) + np.diag(a[1:], -1) + np.diag(c[:-1], 1)
    ref = np.linalg.solve(A, d)
    assert np.abs(x - ref).max() < 1e-9 * max(1, np.abs(ref).max())


def test_pressure_pin_rows():
    o = make(N3=32)
    _, (a, b, c) = o.tri_solve(3, 0, 0, 0, np.ones(32, complex))
    assert b[0] == 1 and c[0] == 0 and a[1] == 0 and b[1] == 1 and c[1] == 0


# ---------------------------------------------------------------- CN amplification
def _set_state(o, **fields):
    for name in ("u1", "u2", "u3", "b", "p"):
        o.set_modal(name, fields.get(name, np.zeros(o.modal_shape, complex)))


@pytest.mark.parametrize("j1,Pr", [(1, 1.0), (3, 0.7)])
def test_crank_nicolson_amplification_closed_form(j1, Pr):
    """A vertically uniform buoyancy mode decays by prod_k (1-a_k)/(1+a_k), a_k = h_k k^2/(2 Re Pr)."""
    N1, N3, Re, dt = 24, 32, 50.0, 0.1
    o = make(N1=N1, N3=N3, Re=Re, Pr=Pr, Ri=0.0)
    o.set_background(np.zeros(N3))
    b = np.zeros(o.modal_shape, complex)
    b[0, j1, :] = 1.0 + 0.5j
    _set_state(o, b=b)
    o.set_dt(dt)
    o.step(1)
    out = o.get_modal("b")[0, j1]
    k2 = (2 * np.pi * j1 / LX) ** 2
    g = 1.0
    for hk in (dt * 8 / 15, dt * 2 / 15, dt * 5 / 15):
        a = 0.5 * hk * k2 / (Re * Pr)
        g *= (1 - a) / (1 + a)
    interior = slice(3, N3 - 3)
    assert np.allclose(out[interior], g * (1.0 + 0.5j), rtol=1e-12)


# ---------------------------------------------------------------- projection
@pytest.mark.parametrize("dim,N2", [(O.TWO_D, 1), (O.THREE_D, 12)])
def test_projection_is_divergence_free(dim, N2):
    N1, N3 = 24, 32
    o = make(N1=N1, N2=N2, N3=N3, dim=dim)
    rng = np.random.default_rng(9)
    shp = o.modal_shape
    fields = {n: (rng.standard_normal(shp) + 1j * rng.standard_normal(shp)) for n in ("u1", "u2", "u3")}
    _set_state(o, **fields)
    o.filter_all()
    o.remove_divergence(1.0)
    u1, u2, u3 = o.get_modal("u1"), o.get_modal("u2"), o.get_modal("u3")
    kx = 2 * np.pi / LX * np.array([j if j <= N1 // 3 else 0 for j in range(N1 // 2 + 1)])
    if dim == O.THREE_D:
        ky = 2 * np.pi / (LX / 2) * np.array([j if j <= N2 // 3 else (j - N2 if j >= N2 - N2 // 3 else 0)
                                              for j in range(N2)])
    else:
        ky = np.zeros(1)
    mask = _mask(N1, N2, dim == O.THREE_D)
    worst = 0.0
    for j2 in range(N2):
        for j1 in range(N1 // 2 + 1):
            if not mask[j2, j1]:
                continue
            dz = o.apply_op("dz_DN", u3[j2, j1].real) + 1j * o.apply_op("dz_DN", u3[j2, j1].imag)
            div = 1j * kx[j1] * u1[j2, j1] + 1j * ky[j2] * u2[j2, j1] * (dim == O.THREE_D) + dz
            worst = max(worst, np.abs(div[2:N3 - 2]).max())
    assert worst < 1e-9


# ---------------------------------------------------------------- InnerProd
def test_inner_prod_parseval_2d():
    N1, N3 = 24, 32
    o = make(N1=N1, N3=N3)
    z, zf, DY, DYF = o.grid()
    rng = np.random.default_rng(2)
    x = rng.standard_normal((1, N1, N3))
    m = o.to_modal(x)
    xf = o.to_nodal(m)  # the filtered field
    w = np.zeros(N3)
    w[1:N3 - 1] = DYF[1:N3 - 1]
    ref = np.sum(w * np.mean(xf[0] ** 2, axis=0))
    assert o.inner_prod(m, m, O.NEUMANN) == pytest.approx(ref, rel=1e-12)


def test_inner_prod_3d_double_counts_j1_zero():
    """Integration.h:124-143 weights every mode but (0,0) by 2, including (0, j2 != 0)."""
    N1, N2, N3 = 12, 6, 16
    o = make(N1=N1, N2=N2, N3=N3, dim=O.THREE_D)
    z, zf, DY, DYF = o.grid()
    rng = np.random.default_rng(4)
    x = rng.standard_normal((N2, N1, N3))
    m = o.to_modal(x)
    xf = o.to_nodal(m)
    w = np.zeros(N3)
    w[1:N3 - 1] = DYF[1:N3 - 1]
    parseval = np.sum(w * np.mean(xf ** 2, axis=(0, 1)))
    extra = np.sum(w * np.sum(np.abs(m[1:, 0, :]) ** 2, axis=0))
    assert o.inner_prod(m, m, O.NEUMANN) == pytest.approx(parseval + extra, rel=1e-12)


# ---------------------------------------------------------------- time stepping
def _kh_state(o, seed=0, amp=1e-2):
    z, zf, DY, DYF = o.grid()
    U = np.tanh(zf)
    U[1] = U[2]; U[0] = U[1]; U[-2] = U[-3]; U[-1] = U[-2]
    o.set_background(U)
    rng = np.random.default_rng(seed)
    for name in ("u1", "u3", "b"):
        zz = z if name == "u3" else zf
        a = np.zeros(o.modal_shape, complex)
        a[0, 1:3, :] = amp * (rng.standard_normal((2, 1)) + 1j * rng.standard_normal((2, 1))) * np.exp(-(zz / 2) ** 2)
        o.set_modal(name, a)
    o.set_modal("u2", np.zeros(o.modal_shape, complex))
    o.set_modal("p", np.zeros(o.modal_shape, complex))
    o.filter_all()
    o.populate_nodal()
    o.remove_divergence(0.0)


def test_temporal_order():
    """From a state that already satisfies the discrete boundary rows, u and b
    converge at second order or better under dt halving, while p converges at
    first order: RemoveDivergence adds q/h_k to p, which the reference itself
    calls "effectively forward euler" (IMEXRK.cpp:69-72)."""
    o = make(N1=24, N3=48, Re=500.0)
    _kh_state(o)
    o.set_dt(0.0005)
    o.step(2)
    start = {n: o.get_modal(n) for n in ("u1", "u2", "u3", "b", "p")}
    U = None
    res = []
    for dt in (0.04, 0.02, 0.01, 0.005):
        q = make(N1=24, N3=48, Re=500.0)
        z, zf, DY, DYF = q.grid()
        U = np.tanh(zf)
        U[1] = U[2]; U[0] = U[1]; U[-2] = U[-3]; U[-1] = U[-2]
        q.set_background(U)
        for n, v in start.items():
            q.set_modal(n, v)
        q.populate_nodal()
        q.set_dt(dt)
        q.step(int(round(0.8 / dt)))
        res.append({n: q.get_modal(n) for n in ("u1", "u3", "b", "p")})
    for n in ("u1", "u3", "b", "p"):
        r = [np.linalg.norm(res[k][n] - res[k + 1][n]) / np.linalg.norm(res[k + 1][n] - res[k + 2][n]) for k in (0, 1)]
        if n == "p":
            assert all(1.8 < x < 2.3 for x in r), (n, r)
        else:
            assert all(x > 3.5 for x in r), (n, r)


def test_inconsistent_start_costs_one_order():
    """Starting from fields that violate the discrete end rows costs O(dt) once."""
    res = []
    for dt in (0.04, 0.02, 0.01):
        o = make(N1=24, N3=48, Re=500.0)
        _kh_state(o)
        o.set_dt(dt)
        o.step(int(round(0.8 / dt)))
        res.append(o.get_modal("u1"))
    ratio = np.linalg.norm(res[0] - res[1]) / np.linalg.norm(res[1] - res[2])
    assert 1.8 < ratio < 2.3, ratio


def test_kelvin_helmholtz_growth_rate():
    """Michalke (1964): tanh shear layer, alpha = 0.4446, sigma ~ 0.19 inviscid; Re = 1000 slightly lower."""
    o = make(N1=16, N3=128, Re=1000.0, Ri=0.0)
    _kh_state(o, amp=1e-6)
    o.set_dt(0.02)
    E = []
    for _ in range(16):
        o.step(100)
        E.append(o.energy()[0])
    E = np.array(E)
    sigma = 0.5 * np.log(E[-1] / E[-2]) / 2.0
    assert 0.17 < sigma < 0.195, sigma


def test_stratification_reduces_growth():
    rates = []
    for Ri in (0.0, 0.1):
        o = make(N1=16, N3=96, Re=1000.0, Ri=Ri)
        _kh_state(o, amp=1e-6)
        o.set_dt(0.02)
        E = []
        for _ in range(12):
            o.step(100)
            E.append(o.energy()[0] + o.energy()[1])
        rates.append(0.5 * np.log(E[-1] / E[-2]) / 2.0)
    assert rates[1] < rates[0]


def test_fp32_oracle_tracks_fp64_oracle():
    out = []
    for prec in (8, 4):
        o = make(N1=48, N3=64, prec=prec)
        _kh_state(o, amp=5e-2)
        o.set_dt(0.01)
        o.step(50)
        out.append(o.energy())
    assert abs(out[1][0] - out[0][0]) / out[0][0] < 1e-4
    assert abs(out[1][1] - out[0][1]) / out[0][1] < 1e-4


# ---------------------------------------------------------------- BuildRHSLinear
@pytest.mark.parametrize("dim,N2", [(O.TWO_D, 1), (O.THREE_D, 12)])
def test_linearised_rhs_is_the_derivative(dim, N2):
    """BuildRHSLinear about w applied to v equals the derivative of BuildRHS at w in
    direction v.  The products are quadratic, so the central difference
    (R(w+ev) - R(w-ev)) / 2e is exact up to rounding."""
    from stratiflow_amd import ic
    N1, N3 = 24, 32
    o0 = make(N1=N1, N2=N2, N3=N3, dim=dim)
    z, zf, _, _ = o0.grid()
    U = ic.tanh_background(zf)
    w = ic.perturbation(N1, N2, N3, dim, z, zf, 5, amplitude=0.2)
    v = ic.perturbation(N1, N2, N3, dim, z, zf, 6, amplitude=0.1)
    rn = ["r1", "r3", "rB"] + (["r2"] if dim == O.THREE_D else [])
    eps = 1e-3

    def r_nonlinear(sign):
        o = make(N1=N1, N2=N2, N3=N3, dim=dim)
        o.set_background(U)
        for n in w:
            o.set_modal(n, w[n] + sign * eps * v[n])
        o.filter_all()
        o.populate_nodal()
        o.stage(0)
        return {n: o.get_modal(n) for n in rn}

    rp, rm = r_nonlinear(+1), r_nonlinear(-1)
    o = make(N1=N1, N2=N2, N3=N3, dim=dim)
    o.set_background(U)
    for n in v:
        o.set_modal(n, v[n])
    names = {"u1": "u1t", "u2": "u2t", "u3": "u3t", "b": "bt"}
    for n in w:
        o.set_modal(names[n], w[n])
    o.set_linear_background()
    o.filter_all()
    o.populate_nodal()
    o.stage(0, linear=True)
    for n in rn:
        fd = (rp[n] - rm[n]) / (2 * eps)
        lin = o.get_modal(n)
        assert np.linalg.norm(fd - lin) / np.linalg.norm(lin) < 1e-9, n


# ---------------------------------------------------------------- EvolveBackground
def _background_stage(U, N3=48, Re=700.0, dt=0.02):
    o = make(N1=12, N3=N3, Re=Re)
    o.set_background(U)
    o.set_evolve_background(True)
    o.set_dt(dt)
    o.stage(0)  # zero state: only U_ changes
    return o, o.background()


def test_evolve_background_constant_profile():
    """A uniform U_ is a steady solution of the interior rows; the Neumannified
    boundary rows (Differentiation.h:17-35) put the ghost values to zero after
    ZeroEnds of the 1D right-hand side (Field.h:625-636)."""
    N3 = 48
    _, U1 = _background_stage(np.full(N3, 0.7), N3)
    assert np.allclose(U1[1:N3 - 1], 0.7, rtol=0, atol=1e-14)
    assert U1[0] == 0.0 and U1[N3 - 1] == 0.0


def test_evolve_background_matches_dense_crank_nicolson():
    """One stage of IMEXRK.cpp:100-104 against a dense solve written from the
    reference's definitions: RU = U + (h/2Re) Dzz U with h = dt*8/15, ZeroEnds
    (slots 0, 1, N3-2, N3-1), then (I - (h/2Re) Dzz) Neumannified (rows 0 and
    N3-1 identity, row 1 = e1 - e2, row N3-2 = e_{N3-3} - e_{N3-2}).  Dzz is the
    oracle's banded operator, pinned on quadratics and as Dz o Dz above."""
    N3, Re, dt = 48, 700.0, 0.02
    o0 = make(N1=12, N3=N3)
    z, zf, _, _ = o0.grid()
    U = np.tanh(zf) + 0.3 * np.exp(-zf ** 2)
    o, U1 = _background_stage(U, N3, Re, dt)
    D = np.stack([o0.apply_op("dzz_N", e) for e in np.eye(N3)], axis=1)
    c = 0.5 * (dt * 8.0 / 15.0) / Re
    rhs = U + c * D @ U
    rhs[[0, 1, N3 - 2, N3 - 1]] = 0
    T = np.eye(N3) - c * D
    T[0, :] = 0
    T[0, 0] = 1
    T[1, :] = 0
    T[1, 1], T[1, 2] = 1, -1
    T[N3 - 2, :] = 0
    T[N3 - 2, N3 - 3], T[N3 - 2, N3 - 2] = 1, -1
    T[N3 - 1, :] = 0
    T[N3 - 1, N3 - 1] = 1
    want = np.linalg.solve(T, rhs)
    assert np.allclose(U1, want, rtol=0, atol=1e-13)
    # and it diffuses: the interior profile moves toward its mean
    assert np.abs(np.diff(U1[1:N3 - 1])).max() < np.abs(np.diff(U[1:N3 - 1])).max()


# ------------------------------------------------------- LinearEvolve dt
@pytest.mark.parametrize("T", [5.0, 37.0])
def test_linear_evolve_time_steps_closed_form(T):
    """LinearEvolve's bookkeeping (StateVector.cpp:133-146) about the zero state:
    CFLlinear is dt*N1/L1 there (IMEXRK.h:376-393), so the first rescale sets
    dt = 0.4*L1/N1 and, every 100 steps, keeps it; once fewer than 100 steps are
    left, dt becomes remaining/(int(remaining/dt) + 1).  The step sequence of
    the restatement must be exactly that."""
    from oracle import krylov
    N1, N3 = 24, 32
    o = make(N1=N1, N3=N3)
    L1 = 2 * np.pi / 0.44
    dts = []
    step = o.step_linear

    def record(n):
        dts.append(o.dt)
        step(n)

    o.step_linear = record
    zero = {n: np.zeros(o.modal_shape, dtype=complex) for n in ("u1", "u3", "b")}
    krylov.linear_evolve(o, dict(zero), dict(zero), T)
    d = 0.4 * L1 / N1
    expect, t = [], 0.0
    while t + 1e-4 < T:
        if len(expect) % 100 == 0:
            dd = d
            left = int((T - t) / dd) + 1
            if left < 100:
                dd = (T - t) / left
        expect.append(dd)
        t += dd
    assert len(dts) == len(expect)
    assert np.allclose(dts, expect, rtol=1e-12, atol=0)
    assert abs(sum(dts) - T) < 1e-9
