"""Dimensionality::TwoAndAHalf (Constants.h:94) resolves the spanwise
direction exactly as ThreeDimensional: GridParams::ThirdDimension() is true for
both (Parameters.h:13-17) and the solver never branches on the value itself;
only StateVector::PlotAll (StateVector.h:495) does.  A TwoAndAHalf context
therefore steps bit for bit as a ThreeDimensional one."""
import numpy as np
import pytest

import stratiflow_amd as sf
from stratiflow_amd import ic
from helpers import LX


def test_dimensionality_codes_follow_the_reference_enum():
    assert (sf.THREE_D, sf.TWO_D, sf.TWO_AND_A_HALF) == (0, 1, 2)
    z = np.linspace(-1, 1, 32)
    assert ic.active_j2(12, sf.TWO_AND_A_HALF) == ic.active_j2(12, sf.THREE_D)
    a = ic.perturbation(24, 12, 32, sf.TWO_AND_A_HALF, z, z, 3)
    b = ic.perturbation(24, 12, 32, sf.THREE_D, z, z, 3)
    assert sorted(a) == sorted(b) == ["b", "u1", "u2", "u3"]
    for n in a:
        assert np.array_equal(a[n], b[n])


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [4, 8])
def test_two_and_a_half_steps_as_three_dimensional(prec):
    runs = []
    for dim in (sf.THREE_D, sf.TWO_AND_A_HALF):
        s = sf.IMEXRK(24, 12, 32, dim, LX, LX / 2, 8.0, 900.0, 0.1, 1.0, precision=prec)
        z, zf = s.vertical_points()
        s.SetBackground(ic.tanh_background(zf), ic.tanh_stratification(zf))
        for n, a in ic.perturbation(24, 12, 32, dim, z, zf, 11, amplitude=0.05).items():
            s.set(n, a)
        s.FilterAll()
        s.PopulateNodalVariables()
        s.RemoveDivergence(0.0)
        s.deltaT = 0.02
        s.TimeStep(4)
        runs.append({n: s.get(n) for n in ("u1", "u2", "u3", "b", "p")})
        s.close()
    for n in runs[0]:
        assert np.array_equal(runs[0][n], runs[1][n]), n
