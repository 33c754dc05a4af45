"""FlowParams::EvolveBackground: the background U_ diffuses by one
Crank-Nicolson solve per stage (IMEXRK.cpp:100-104) and the products of the
next stage see the new U_.  The GPU path (the (0,0) wave of the stage kernel)
against the oracle (pinned by a dense solve in test_oracle.py), 2D and 3D."""
import numpy as np
import pytest

from helpers import LX, fields, rel_l2

pytestmark = pytest.mark.gpu


def _pair(N1, N2, N3, dim, prec, seed=5, Re=800.0, via_reset=False):
    import oracle
    import stratiflow_amd as sf
    from stratiflow_amd import ic
    s = sf.IMEXRK(N1, N2, N3, dim, LX, LX / 2, 10.0, Re, 0.12, 1.0, precision=prec, evolve_background=not via_reset)
    if via_reset:  # switched on mid-run through FlowParams (sf_set_flow)
        s.ResetForParams(EvolveBackground=True)
    o = oracle.Oracle(N1, N2, N3, dim, LX, LX / 2, 10.0, Re, 0.12, 1.0, precision=8)
    o.set_evolve_background(True)
    z, zf = s.vertical_points()
    U, S = ic.tanh_background(zf), ic.tanh_stratification(zf)
    s.SetBackground(U, S)
    o.set_background(U, S)
    for n, a in ic.perturbation(N1, N2, N3, dim, z, zf, seed, amplitude=0.05).items():
        s.set(n, a)
        o.set_modal(n, a)
    s.FilterAll()
    s.PopulateNodalVariables()
    s.RemoveDivergence(0.0)
    o.filter_all()
    o.populate_nodal()
    o.remove_divergence(0.0)
    s.deltaT = 0.02
    o.set_dt(0.02)
    return s, o, U


@pytest.mark.parametrize("shape", [(48, 1, 64, 1), (24, 12, 32, 0)])
@pytest.mark.parametrize("prec,tol", [(8, 1e-10), (4, 2e-4)])
def test_evolve_background_steps_match_oracle(shape, prec, tol):
    N1, N2, N3, dim = shape
    s, o, U0 = _pair(N1, N2, N3, dim, prec)
    s.TimeStep(3)
    o.step(3)
    Ug, Uo = s.Background(), o.background()
    eu = np.abs(Ug - Uo).max() / np.abs(Uo).max()
    print("%s fp%d: U_ moved %.3e, gpu vs oracle %.3e" % (shape, 8 * prec, np.abs(Uo - U0).max(), eu), flush=True)
    assert np.abs(Uo - U0).max() > 1e-4  # it did diffuse
    assert eu < (1e-13 if prec == 8 else 1e-6)
    for f in fields(dim):
        e = rel_l2(s.get(f), o.get_modal(f))
        print("  %s %.3e" % (f, e), flush=True)
        assert e < (tol if f != "p" else 10 * tol), f
    # a linearised step also evolves U_ (CrankNicolson checks the flag itself)
    names = ["u1", "u2", "u3", "b"] if dim == 0 else ["u1", "u3", "b"]
    st = {n: s.state(n) for n in names}
    s.SetLinearBackground(st["u1"], st.get("u2"), st["u3"], st["b"])
    for n in names:
        o.set_modal(n + "t", o.get_modal(n))
    o.set_linear_background()
    s.TimeStepLinear(1)
    o.step_linear(1)
    assert np.abs(s.Background() - o.background()).max() / np.abs(Uo).max() < (1e-13 if prec == 8 else 1e-6)
    for f in fields(dim):
        e = rel_l2(s.get(f), o.get_modal(f))
        assert e < (tol if f != "p" else 10 * tol), f
    s.close()


def test_evolve_background_switched_on_through_flow_params():
    """EvolveBackground set after construction (IMEXRK::ResetForParams with
    the new FlowParams) diffuses U_ exactly as when it is set at construction."""
    s, o, U0 = _pair(48, 1, 64, 1, 8, via_reset=True)
    s.TimeStep(2)
    o.step(2)
    Uo = o.background()
    assert np.abs(Uo - U0).max() > 1e-5
    assert np.abs(s.Background() - Uo).max() / np.abs(Uo).max() < 1e-13
    for f in fields(1):
        assert rel_l2(s.get(f), o.get_modal(f)) < (1e-10 if f != "p" else 1e-9), f
    s.close()
