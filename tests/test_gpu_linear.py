"""TimeStepLinear (IMEXRK.cpp:26-41, BuildRHSLinear :195-258) on the GPU vs the oracle."""
import numpy as np
import pytest

import stratiflow_amd as sf
from stratiflow_amd import ic
from helpers import fields, make_pair, rel_l2

pytestmark = pytest.mark.gpu


def _with_base(shape, prec, oracle_prec=8):
    N1, N2, N3, dim = shape
    s, o = make_pair(N1, N2, N3, dim, gpu_prec=prec, oracle_prec=oracle_prec, amplitude=0.01, project=False)
    z, zf = s.vertical_points()
    base = ic.perturbation(N1, N2, N3, dim, z, zf, 99, amplitude=0.2)
    names = {"u1": "u1t", "u2": "u2t", "u3": "u3t", "b": "bt"}
    gpu_fields = {}
    for n, a in base.items():
        f = s.field(sf.MODAL, sf.DIRICHLET if n == "u3" else sf.NEUMANN)
        f.upload(a)
        gpu_fields[n] = f
        o.set_modal(names[n], a)
    s.SetLinearBackground(gpu_fields["u1"], gpu_fields.get("u2"), gpu_fields["u3"], gpu_fields["b"])
    o.set_linear_background()
    # LinearEvolve's start-up (StateVector.cpp:109-115)
    s.FilterAll()
    s.PopulateNodalVariables()
    s.RemoveDivergence(0.0)
    o.filter_all()
    o.populate_nodal()
    o.remove_divergence(0.0)
    return s, o, gpu_fields


@pytest.mark.parametrize("shape", [(48, 1, 64, 1), (24, 12, 32, 0)])
def test_linear_steps_fp64(shape):
    s, o, keep = _with_base(shape, 8)
    s.TimeStepLinear(5)
    o.step_linear(5)
    for f in fields(shape[3]):
        e = rel_l2(s.get(f), o.get_modal(f))
        print("linear fp64", f, e)
        assert e < 1e-9, f


@pytest.mark.parametrize("shape", [(48, 1, 64, 1), (24, 12, 32, 0)])
def test_linear_steps_fp32(shape):
    s, o, keep = _with_base(shape, 4)
    s.TimeStepLinear(10)
    o.step_linear(10)
    for f in fields(shape[3]):
        if f == "p":
            continue
        e = rel_l2(s.get(f), o.get_modal(f))
        print("linear fp32", f, e)
        assert e < 1e-4, f
