"""Error behaviour of the C ABI on a live context: invalid arguments come back
as SF_ERR_ARG (1) or SF_ERR_UNSUPPORTED (3) with a message naming the rule,
nothing is launched, and the context stays usable afterwards.  Every case here
is rejected by the host before any kernel runs."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_invalid_arguments_are_reported_and_recoverable():
    import stratiflow_amd as sf
    s = sf.IMEXRK(48, 1, 64, sf.TWO_D, 14.28, 1.0, 10.0, 1000.0, 0.1, 1.0, precision=8)
    n = s.field(sf.NODAL, sf.NEUMANN)
    m = s.field(sf.MODAL, sf.DIRICHLET)
    with pytest.raises(sf.SFError) as e:
        n.ToModal(m)  # ToModal keeps the boundary condition (Field.h:736)
    assert e.value.code == 1 and "ToModal" in str(e.value)
    with pytest.raises(sf.SFError) as e:
        s.deltaT = 0.0
    assert e.value.code == 1 and "dt" in str(e.value)
    with pytest.raises(sf.SFError) as e:
        s.InnerProd(s.state("u1"), s.state("u3"))  # Neumann against Dirichlet
    assert e.value.code == 1 and "InnerProd" in str(e.value)
    # the solver's own fields cannot stand for the direct state of the adjoint step
    st = [s.state(k) for k in ("u1", "u2", "u3", "b")]
    with pytest.raises(sf.SFError) as e:
        s.TimeStepAdjoint(st, st)
    assert e.value.code == 1 and "own" in str(e.value)
    with pytest.raises(sf.SFError) as e:
        s.Stage(3)
    assert e.value.code == 1
    net = sf.Net(s, hidden=64, layers=4, seed=1)
    with pytest.raises(sf.SFError) as e:
        net.set_dilations([1, 2, 4, 9])
    assert e.value.code == 1 and "dilation" in str(e.value)
    with pytest.raises(ValueError):
        net.set_dilations([1, 2])
    net.close()
    # the context still steps, and the rejected dt left the old one in place
    s.deltaT = 0.01
    s.TimeStep(2)
    assert s.deltaT == 0.01
    assert np.isfinite(s.energies()[0])
    s.close()


@pytest.mark.parametrize("shape,code", [((48, 2, 64, 1), 1),    # 2D needs N2 == 1 (IMEXRK.h:43)
                                        ((50, 1, 64, 1), 3),    # FFT sizes are 2^a 3^b
                                        ((48, 1, 62, 1), 1),    # N3 % 4 == 0
                                        ((48, 5, 64, 0), 3)])   # 3D needs N2 >= 6
def test_unsupported_grids(shape, code):
    import stratiflow_amd as sf
    N1, N2, N3, dim = shape
    with pytest.raises(sf.SFError) as e:
        sf.IMEXRK(N1, N2, N3, dim, 14.28, 7.14, 10.0, 1000.0, 0.1, 1.0)
    assert e.value.code == code


def test_closing_a_solver_frees_its_fields_and_networks_first():
    """sf_ctx_destroy requires the context's fields and networks to be freed
    first; the Python solver does that for the objects it made, so dropping
    them after close() is harmless."""
    import gc
    import stratiflow_amd as sf
    s = sf.IMEXRK(48, 1, 64, sf.TWO_D, 14.28, 1.0, 10.0, 1000.0, 0.1, 1.0)
    f = s.field(sf.MODAL, sf.NEUMANN)
    net = sf.Net(s, hidden=32, layers=2, seed=0)
    s.close()
    assert f.h is None and net.h is None
    del f, net
    gc.collect()
