"""sf_cfl_linear_noalloc (scratch in a static device array) against sf_cfl_linear.

Both entry points run the same kernels; the no-allocation one takes its
nodal scratch field and reduction partials from a fixed per-GPU array.  The
results must be identical bit for bit, on two calls in a row of which the
second has the smaller grid, then again on the first grid (the array keeps
the last call's contents), in 2D and 3D, and a grid whose nodal field does
not fit the array is rejected with SF_ERR_ARG.
"""
import ctypes

import numpy as np
import pytest

from helpers import make_pair

pytestmark = pytest.mark.gpu


def _linearised(N1, N2, N3, dim, seed, prec=4):
    s, _ = make_pair(N1, N2, N3, dim, gpu_prec=prec, oracle_prec=8, amplitude=0.3, seed=seed)
    u2 = s.state("u2") if dim == 0 else None
    s.SetLinearBackground(s.state("u1"), u2, s.state("u3"), s.state("b"))
    return s


def _both(s):
    import stratiflow_amd as sf
    a, b = ctypes.c_double(), ctypes.c_double()
    assert sf.lib().sf_cfl_linear(s.h, ctypes.byref(a)) == 0
    rc = sf.lib().sf_cfl_linear_noalloc(s.h, ctypes.byref(b))
    assert rc == 0, sf.lib().sf_last_error(s.h).decode()
    print("cfl_linear %.17e noalloc %.17e" % (a.value, b.value), flush=True)
    return a.value, b.value


@pytest.mark.parametrize("big,small", [((48, 1, 64, 1), (24, 1, 32, 1)), ((48, 24, 64, 0), (24, 12, 32, 0))])
def test_noalloc_equals_allocating_entry_point(big, small):
    s1 = _linearised(*big, seed=31)
    s2 = _linearised(*small, seed=32)
    a1, b1 = _both(s1)
    a2, b2 = _both(s2)  # second call, smaller input
    assert a1 == b1 and np.isfinite(a1) and a1 > 0
    assert a2 == b2 and np.isfinite(a2) and a2 > 0
    assert a1 != a2
    a3, b3 = _both(s1)  # the array holds the smaller call's data now
    assert a3 == b3 == a1
    s1.close()
    s2.close()


def test_noalloc_fp64():
    s = _linearised(48, 1, 64, 1, seed=33, prec=8)
    a, b = _both(s)
    assert a == b
    s.close()


def test_noalloc_rejects_grids_beyond_its_scratch():
    import stratiflow_amd as sf
    # 256*128*256 FP32 values are 32 MiB, more than the array's nodal part
    s = sf.IMEXRK(256, 128, 256, 0, 2 * np.pi / 0.44, np.pi / 0.44, 10.0, 1000.0, 0.1, 1.0, precision=4)
    out = ctypes.c_double(-1.0)
    rc = sf.lib().sf_cfl_linear_noalloc(s.h, ctypes.byref(out))
    msg = sf.lib().sf_last_error(s.h).decode()
    print("rejected:", rc, msg, flush=True)
    assert rc == 1  # SF_ERR_ARG
    assert "exceeds the scratch" in msg
    assert out.value == -1.0
    s.close()
