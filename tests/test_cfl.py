"""CFL and CFLlinear (IMEXRK.h:355-393) of the GPU path against the oracle.

The oracle's CFL functions are pinned by the closed forms of uniform flows
(test_oracle.py); here the GPU reductions (nodal max-abs with the background
added, the grid spacings, the dt factor) are compared with them on seeded
perturbed states, 2D and 3D, FP32 and FP64, and the dt update of the Python
mirror's CFL()/CFLlinear() (targets 0.8 and 0.4) is checked.
"""
import numpy as np
import pytest

from helpers import make_pair

pytestmark = pytest.mark.gpu

SHAPES = [(48, 1, 64, 1), (24, 12, 32, 0), (64, 1, 100, 1)]


@pytest.mark.parametrize("prec,tol", [(8, 1e-12), (4, 1e-6)])
@pytest.mark.parametrize("shape", SHAPES)
def test_cfl_matches_oracle(shape, prec, tol):
    N1, N2, N3, dim = shape
    s, o = make_pair(N1, N2, N3, dim, gpu_prec=prec, oracle_prec=prec, amplitude=0.3, dt=0.013, seed=7)
    # FullEvolve's start-up left the populated (pre-projection) nodal fields
    # in place; CFL reads those, as the reference does at step 0
    g0 = s.cfl_number()
    o0 = o.cfl()
    print("cfl start %s: gpu %.9e oracle %.9e" % (shape, g0, o0), flush=True)
    assert abs(g0 - o0) / o0 < tol
    # after two steps the nodal fields are those of the last stage
    s.TimeStep(2)
    o.step(2)
    g1 = s.cfl_number()
    o1 = o.cfl()
    print("cfl after 2 steps %s: gpu %.9e oracle %.9e" % (shape, g1, o1), flush=True)
    assert abs(g1 - o1) / o1 < tol
    # IMEXRK::CFL rescales dt to the target CFL 0.8
    dt0 = s.deltaT
    c = s.CFL()
    assert c == pytest.approx(g1, rel=1e-12)
    assert s.deltaT == pytest.approx(dt0 * 0.8 / c, rel=1e-6)
    assert s.cfl_number() == pytest.approx(0.8, rel=1e-5)


@pytest.mark.parametrize("prec,tol", [(8, 1e-12), (4, 1e-6)])
@pytest.mark.parametrize("shape", SHAPES)
def test_cfl_linear_matches_oracle(shape, prec, tol):
    N1, N2, N3, dim = shape
    s, o = make_pair(N1, N2, N3, dim, gpu_prec=prec, oracle_prec=prec, amplitude=0.4, dt=0.01, seed=8)
    # linearise about the projected state (SetBackground(u1..b), IMEXRK.h:334-352)
    names = ["u1", "u2", "u3", "b"] if dim == 0 else ["u1", "u3", "b"]
    st = {n: s.state(n) for n in names}
    s.SetLinearBackground(st["u1"], st.get("u2"), st["u3"], st["b"])
    for n in names:
        o.set_modal(n + "t", o.get_modal(n))
    o.set_linear_background()
    g = s.cfl_number(linear=True)
    r = o.cfl(linear=True)
    print("cfl_linear %s: gpu %.9e oracle %.9e" % (shape, g, r), flush=True)
    assert abs(g - r) / r < tol
    dt0 = s.deltaT
    c = s.CFLlinear()
    assert c == pytest.approx(g, rel=1e-12)
    assert s.deltaT == pytest.approx(dt0 * 0.4 / c, rel=1e-6)
