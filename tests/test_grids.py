"""GPU path vs oracle at the edges of the supported grids: vertical columns
that do not fill whole wavefronts (N3 below 64 or not a multiple of 64), the
smallest 3D grid (N2 = 6), the 256x1x256 Newton-search grid, the largest
sizes (N1 = N3 = 512), and N1 = 192, whose x FFT (16 x 12) runs its second
pass on lane pairs of radix 6 (the other N1 cover radix 2, 3, 4, 8, 12 and 16
per lane, and N1 = 48 the unsplit radix-3 pass).  FP64 contexts support N1 <= 384, so the largest grid
runs in FP32 against the FP32 and FP64 oracles."""
import numpy as np
import pytest

from helpers import fields, make_pair, rel_l2

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape", [(256, 1, 256, 1), (64, 1, 100, 1), (48, 1, 512, 1), (12, 6, 16, 0),
                                   (96, 48, 36, 0), (96, 48, 100, 0), (192, 1, 64, 1), (192, 12, 36, 0)])
def test_edge_grids_fp64(shape):
    """Three full steps in FP64: every prognostic field to 1e-9 (relative L2)."""
    N1, N2, N3, dim = shape
    print("grid", shape, flush=True)
    s, o = make_pair(N1, N2, N3, dim, gpu_prec=8, amplitude=0.05)
    s.TimeStep(3)
    o.step(3)
    for f in fields(dim):
        e = rel_l2(s.get(f), o.get_modal(f))
        print("  %s %.3e" % (f, e), flush=True)
        assert e < 1e-9, (f, e)
    s.close()


@pytest.mark.parametrize("shape", [(512, 1, 512, 1), (96, 48, 100, 0)])
def test_edge_grids_fp32(shape):
    """Ten FP32 steps: KE and PE within 1e-4 of the FP32 oracle, every field
    within 1e-4 of the FP64 oracle and within 10x the FP32 oracle's own error."""
    N1, N2, N3, dim = shape
    print("grid", shape, flush=True)
    s, o32 = make_pair(N1, N2, N3, dim, gpu_prec=4, oracle_prec=4, amplitude=0.05)
    _, o64 = make_pair(N1, N2, N3, dim, gpu_prec=4, oracle_prec=8, amplitude=0.05)
    s.TimeStep(10)
    o32.step(10)
    o64.step(10)
    ke, pe = s.energies()
    ko, po = o32.energy()
    assert abs(ke - ko) / ko < 1e-4 and abs(pe - po) / po < 1e-4, (ke, ko, pe, po)
    for f in fields(dim):
        eg = rel_l2(s.get(f), o64.get_modal(f))
        er = rel_l2(o32.get_modal(f), o64.get_modal(f))
        print("  %s gpu %.3e ref32 %.3e" % (f, eg, er), flush=True)
        assert eg < 1e-4, (f, eg)
        assert eg <= max(10 * er, 1e-6), (f, eg, er)
    s.close()
