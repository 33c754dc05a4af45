"""Adjoint time stepping: TimeStepAdjoint / BuildRHSAdjoint / UpdateAdjointVariables
(IMEXRK.h:86-126, 183-199, 461-500; IMEXRK.cpp:260-379).

CPU pins of the oracle's restatement, none of which re-types its formulas:
  * transposition: with no background stratification, the adjoint's explicit
    terms (forcing removed) are the transpose of the linearised explicit terms
    under the InnerProd weights, <J u, v> = <u, A v> for divergence-free u, v,
    up to the vertical truncation error, which falls at second order;
  * x<->y swap: in 3D with L1 = L2, N1 = N2 and no background velocity, one
    adjoint step commutes with exchanging the horizontal axes (u1 <-> u2);
  * y-invariant fields: a 3D step on fields that do not depend on y equals the
    2D step (the 3D-only terms vanish and the rest must agree);
  * zero direct state: the step reduces to the linearised step about zero with
    Ri -> 0 and dB/dz -> -Ri (the adjoint's only coupling is Ri*Idir(u3) into b);
  * zero adjoint state: the explicit terms are the forcing 0.5*(b_tot - <b_tot>)
    in r3 and 0.5*(u3_tot - <u3_tot>) in rB, and nothing else.
GPU: sf_step_adjoint against the oracle on identical inputs (FP64 to 1e-12,
FP32 to 1e-4 relative), over several steps with a changing direct state.
"""
import numpy as np
import pytest

import oracle as O

LX = 2 * np.pi / 0.44
NAMES2 = ["u1", "u3", "b"]
NAMES3 = ["u1", "u2", "u3", "b"]
RN = {"u1": "r1", "u2": "r2", "u3": "r3", "b": "rB"}
BC = {"u1": 0, "u2": 0, "u3": 1, "b": 0}


def _oracle(N1, N2, N3, dim, L1=LX, L2=LX / 2, Ri=0.3, Re=1000.0, prec=8):
    return O.Oracle(N1, N2, N3, dim, L1, L2, 10.0, Re, Ri, 1.0, precision=prec)


def _projected(N1, N2, N3, dim, seed, amp, U=None, S=None, L1=LX, L2=LX / 2):
    """Filtered, divergence-free fields from seeded smooth perturbations."""
    from stratiflow_amd import ic
    q = _oracle(N1, N2, N3, dim, L1, L2)
    z, zf, _, _ = q.grid()
    if U is not None:
        q.set_background(U, S)
    pert = ic.perturbation(N1, N2, N3, dim, z, zf, seed, amplitude=amp)
    for n, a in pert.items():
        q.set_modal(n, a)
    q.filter_all()
    q.populate_nodal()
    q.remove_divergence(0.0)
    return {n: q.get_modal(n) for n in (NAMES3 if dim == O.THREE_D else NAMES2)}


def _transpose_mismatch(N1, N2, N3, dim, seeds=(5, 6, 7)):
    from stratiflow_amd import ic
    o = _oracle(N1, N2, N3, dim)
    z, zf, _, _ = o.grid()
    U = ic.tanh_background(zf)
    S0 = np.zeros(N3)
    w = _projected(N1, N2, N3, dim, seeds[0], 0.2, U, S0)
    u = _projected(N1, N2, N3, dim, seeds[1], 0.1, U, S0)
    v = _projected(N1, N2, N3, dim, seeds[2], 0.1, U, S0)
    # J u: explicit terms of the linearised equations about w + U
    d = _oracle(N1, N2, N3, dim)
    d.set_background(U, S0)
    tn = {"u1": "u1t", "u2": "u2t", "u3": "u3t", "b": "bt"}
    for n in u:
        d.set_modal(n, u[n])
        d.set_modal(tn[n], w[n])
    d.set_linear_background()
    d.populate_nodal()
    d.rhs_linear()
    Ju = {n: d.get_modal(RN[n]) for n in u}

    def adjoint_rhs(state):
        a = _oracle(N1, N2, N3, dim)
        a.set_background(U, S0)
        for n in state:
            a.set_modal(n, state[n])
        a.populate_nodal()
        a.rhs_adjoint(w, w, 0.3)
        return {n: a.get_modal(RN[n]) for n in state}

    Av = adjoint_rhs(v)
    Af = adjoint_rhs({n: 0 * v[n] for n in v})
    Av = {n: Av[n] - Af[n] for n in v}

    def ip(a, b):
        return sum(o.inner_prod(a[n], b[n], BC[n]) for n in a)

    return (ip(Ju, v) - ip(u, Av)) / np.sqrt(ip(Ju, Ju) * ip(v, v))


def test_adjoint_terms_transpose_linearised_terms_2d():
    e32 = abs(_transpose_mismatch(48, 1, 32, O.TWO_D))
    e64 = abs(_transpose_mismatch(48, 1, 64, O.TWO_D))
    e128 = abs(_transpose_mismatch(48, 1, 128, O.TWO_D))
    print("transposition mismatch N3=32,64,128: %.3e %.3e %.3e" % (e32, e64, e128))
    # second order in the vertical spacing; a wrong sign, operand or dropped
    # term of the adjoint leaves a mismatch above 5e-5 at N3 = 64
    assert e64 < 1e-5 and e128 < 2.5e-6
    assert e32 / e64 > 3.0 and e64 / e128 > 3.0


def test_adjoint_terms_transpose_linearised_terms_3d():
    # the horizontal products alias differently on the two sides, which leaves
    # a resolution-independent mismatch of order 1e-4 in 3D
    e = abs(_transpose_mismatch(24, 12, 64, O.THREE_D))
    print("3D transposition mismatch %.3e" % e)
    assert e < 5e-4


def _direct_states(N1, N2, N3, dim, nst, seed=11, L1=LX, L2=LX / 2, U=None, dt=0.02):
    """nst consecutive direct states (the intermediate states of AdjointEvolve)."""
    o = _oracle(N1, N2, N3, dim, L1, L2)
    z, zf, _, _ = o.grid()
    if U is not None:
        o.set_background(U)
    w = _projected(N1, N2, N3, dim, seed, 0.2, L1=L1, L2=L2)
    for n in w:
        o.set_modal(n, w[n])
    o.set_dt(dt)
    names = NAMES3 if dim == O.THREE_D else NAMES2
    out = []
    for _ in range(nst):
        out.append({n: o.get_modal(n) for n in names})
        o.step(1)
    return out


def _adjoint_run(o, a0, states, nsteps):
    """AdjointEvolve (StateVector.cpp:158-187) on the oracle o from a0."""
    for n in a0:
        o.set_modal(n, a0[n])
    o.filter_all()
    o.populate_nodal()
    o.remove_divergence(0.0)
    o.prepare_adjoint()
    for s in range(nsteps):
        o.step_adjoint(states[nsteps - 1 - s], states[nsteps - s])
    return {n: o.get_modal(n) for n in list(a0) + ["p"]}


def _swap_xy(o, f):
    """g(x, y) = f(y, x) for a modal field of a 3D oracle with N1 == N2."""
    return o.to_modal(np.ascontiguousarray(np.swapaxes(o.to_nodal(f), 0, 1)))


def test_adjoint_step_commutes_with_xy_swap():
    N, N3 = 12, 16
    L = LX
    states = _direct_states(N, N, N3, O.THREE_D, 3, L1=L, L2=L, dt=0.05)
    a0 = _projected(N, N, N3, O.THREE_D, 21, 0.1, L1=L, L2=L)
    o1 = _oracle(N, N, N3, O.THREE_D, L, L)
    o1.set_dt(0.05)
    r1 = _adjoint_run(o1, a0, states, 2)
    o2 = _oracle(N, N, N3, O.THREE_D, L, L)
    o2.set_dt(0.05)
    sw = {"u1": "u2", "u2": "u1", "u3": "u3", "b": "b"}

    def swap_state(st):
        return {sw[n]: _swap_xy(o2, st[n]) for n in st}

    r2 = _adjoint_run(o2, swap_state(a0), [swap_state(s) for s in states], 2)
    for n in NAMES3 + ["p"]:
        want = _swap_xy(o2, r1[sw.get(n, n)])
        err = np.abs(r2[n] - want).max() / np.abs(want).max()
        assert err < 1e-11, (n, err)


def test_adjoint_step_3d_reduces_to_2d_for_y_invariant_fields():
    from stratiflow_amd import ic
    N1, N2, N3 = 24, 12, 32
    o2 = _oracle(N1, 1, N3, O.TWO_D)
    z, zf, _, _ = o2.grid()
    U = ic.tanh_background(zf)
    o2.set_background(U)
    o2.set_dt(0.02)
    states2 = _direct_states(N1, 1, N3, O.TWO_D, 3, U=U)
    a2 = _projected(N1, 1, N3, O.TWO_D, 22, 0.1)
    r2 = _adjoint_run(o2, a2, states2, 2)

    def lift(st):
        out = {}
        for n in NAMES3:
            m = np.zeros((N2, N1 // 2 + 1, N3), complex)
            if n in st:
                m[0] = st[n][0]
            out[n] = m
        return out

    o3 = _oracle(N1, N2, N3, O.THREE_D)
    o3.set_background(U)
    o3.set_dt(0.02)
    r3 = _adjoint_run(o3, lift(a2), [lift(s) for s in states2], 2)
    for n in NAMES2 + ["p"]:
        err = np.abs(r3[n][0] - r2[n][0]).max() / np.abs(r2[n][0]).max()
        assert err < 1e-11, (n, err)
        assert np.abs(r3[n][1:]).max() < 1e-14 * np.abs(r2[n]).max()
    assert np.abs(r3["u2"]).max() == 0.0


@pytest.mark.parametrize("dim,N2", [(O.TWO_D, 1), (O.THREE_D, 12)])
def test_zero_direct_state_reduces_to_linearised_step(dim, N2):
    N1, N3, Ri = 24, 32, 0.3
    a0 = _projected(N1, N2, N3, dim, 23, 0.1)
    zero = {n: 0 * a0[n] for n in a0}
    adj = _oracle(N1, N2, N3, dim, Ri=Ri)
    adj.set_dt(0.03)
    ra = _adjoint_run(adj, a0, [zero] * 3, 2)
    lin = _oracle(N1, N2, N3, dim, Ri=0.0)
    lin.set_background(np.zeros(N3), -Ri * np.ones(N3))
    lin.set_dt(0.03)
    for n in a0:
        lin.set_modal(n, a0[n])
    lin.filter_all()
    lin.populate_nodal()
    lin.remove_divergence(0.0)
    lin.set_modal("p", np.zeros(lin.modal_shape, complex))
    lin.populate_nodal()
    lin.step_linear(2)
    assert np.abs(ra["b"] - a0["b"]).max() > 1e-5  # the coupling acted
    for n in list(a0) + ["p"]:
        want = lin.get_modal(n)
        err = np.abs(ra[n] - want).max() / np.abs(want).max()
        assert err < 1e-11, (n, err)


@pytest.mark.parametrize("dim,N2", [(O.TWO_D, 1), (O.THREE_D, 12)])
def test_zero_adjoint_state_gives_the_forcing(dim, N2):
    N1, N3 = 24, 32
    w = _projected(N1, N2, N3, dim, 24, 0.2)
    w2 = _projected(N1, N2, N3, dim, 25, 0.2)
    o = _oracle(N1, N2, N3, dim)
    o.populate_nodal()  # zero adjoint state
    f = 0.25
    o.rhs_adjoint(w, w2, f)
    tot = {n: (1 - f) * w2[n] + f * w[n] for n in w}

    def minus_mean(a):
        b = a.copy()
        b[0, 0, :] -= a[0, 0, :].real
        return b

    r3, rB = o.get_modal("r3"), o.get_modal("rB")
    assert np.abs(r3 - 0.5 * minus_mean(tot["b"])).max() < 1e-14
    assert np.abs(rB - 0.5 * minus_mean(tot["u3"])).max() < 1e-14
    assert np.abs(o.get_modal("r1")).max() < 1e-16
    if dim == O.THREE_D:
        assert np.abs(o.get_modal("r2")).max() < 1e-16


# ---------------------------------------------------------------- GPU parity
def _gpu_adjoint(N1, N2, N3, dim, prec, nsteps, dt=0.02):
    import stratiflow_amd as sf
    from stratiflow_amd import ic
    names = NAMES3 if dim == O.THREE_D else NAMES2
    o = _oracle(N1, N2, N3, dim)
    z, zf, _, _ = o.grid()
    U = ic.tanh_background(zf)
    states = _direct_states(N1, N2, N3, dim, nsteps + 1, U=U, dt=dt)
    a0 = _projected(N1, N2, N3, dim, 31, 0.1)
    o.set_background(U)
    o.set_dt(dt)
    want = _adjoint_run(o, a0, states, nsteps)

    s = sf.IMEXRK(N1, N2, N3, dim, LX, LX / 2, 10.0, 1000.0, 0.3, 1.0, precision=prec)
    s.SetBackground(U)
    s.deltaT = dt
    for n in names:
        s.set(n, a0[n])
    s.PrepareRunAdjoint()
    dev = []
    for st in states:
        fs = []
        for n in ("u1", "u2", "u3", "b"):
            if n not in st:
                fs.append(None)
                continue
            f = s.field(sf.MODAL, sf.DIRICHLET if n == "u3" else sf.NEUMANN)
            f.upload(st[n])
            fs.append(f)
        dev.append(fs)
    for k in range(nsteps):
        s.TimeStepAdjoint(dev[nsteps - 1 - k], dev[nsteps - k])
    got = {n: s.get(n) for n in names + ["p"]}
    return got, want


@pytest.mark.gpu
@pytest.mark.parametrize("dim,N1,N2,N3", [(O.TWO_D, 48, 1, 64), (O.THREE_D, 24, 12, 32), (O.TWO_D, 64, 1, 100),
                                          (O.THREE_D, 12, 6, 36)])
@pytest.mark.parametrize("prec,tol", [(8, 1e-12), (4, 2e-4)])
def test_gpu_adjoint_step_matches_oracle(dim, N1, N2, N3, prec, tol):
    print("adjoint parity dim=%d %dx%dx%d prec=%d" % (dim, N1, N2, N3, prec), flush=True)
    got, want = _gpu_adjoint(N1, N2, N3, dim, prec, 3)
    for n in want:
        scale = np.abs(want[n]).max()
        err = np.abs(got[n] - want[n]).max() / scale
        print("  %s max rel err %.3e (scale %.3e)" % (n, err, scale))
        assert err < tol, (n, err)
