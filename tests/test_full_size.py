"""Parity at the throughput configuration, 384 x 192 x 384 (3D), through the C ABI.

* one full TimeStep of the seeded workload (Re=4000, Ri=0.16, dt=0.005, tanh U and
  tanh background buoyancy, as bench.py times it) against
  the FP64 oracle on the same inputs: FP64 to 1e-10, FP32 to 2e-5 relative;
* size-independent properties: ToModal(ToNodal(u)) = u on filtered spectra,
  and the projection leaves a discrete divergence at rounding level.
The oracle step takes seconds on the host's cores at this size.
"""
import time

import numpy as np
import pytest

from helpers import make_pair, rel_l2

pytestmark = pytest.mark.gpu

C2 = (384, 192, 384, 0)


def _pair(prec):
    import oracle
    oracle.set_threads(16)
    N1, N2, N3, dim = C2
    from stratiflow_amd import ic
    # the throughput workload: tanh U and tanh background buoyancy
    return make_pair(N1, N2, N3, dim, Re=4000.0, Ri=0.16, dt=0.005, seed=20261202, gpu_prec=prec, amplitude=1e-3,
                     stratification=ic.tanh_stratification)


def test_full_size_step_fp64_matches_oracle():
    print("full-size step, FP64", flush=True)
    t0 = time.time()
    s, o = _pair(8)
    s.TimeStep(1)
    t1 = time.time()
    o.step(1)
    print("  set-up and GPU step %.1f s, oracle step %.1f s" % (t1 - t0, time.time() - t1), flush=True)
    for n in ("u1", "u2", "u3", "b", "p"):
        err = rel_l2(s.get(n), o.get_modal(n))
        print("  %s rel l2 %.3e" % (n, err), flush=True)
        assert err < (1e-10 if n != "p" else 1e-9), n
    ke, pe = s.energies()
    oke, ope = o.energy()
    assert abs(ke - oke) <= 1e-10 * oke and abs(pe - ope) <= 1e-10 * ope


def test_full_size_step_fp32_within_fp32_rounding():
    """FP32 GPU vs the FP64 oracle, bounded by 3x the FP32 oracle's own error."""
    import oracle
    print("full-size step, FP32", flush=True)
    s, o = _pair(4)
    N1, N2, N3, dim = C2
    # the FP32 oracle starts from the FP64 oracle's projected state and
    # populated nodal fields, exactly what the first stage reads
    o32 = oracle.Oracle(N1, N2, N3, dim, 2 * np.pi / 0.44, np.pi / 0.44, 10.0, 4000.0, 0.16, 1.0, precision=4)
    from stratiflow_amd import ic
    z, zf, _, _ = o32.grid()
    o32.set_background(ic.tanh_background(zf), ic.tanh_stratification(zf))
    for n in ("u1", "u2", "u3", "b", "p"):
        o32.set_modal(n, o.get_modal(n))
    for n in ("U1", "U2", "U3", "B"):
        o32.set_nodal(n, o.get_nodal(n))
    o32.set_dt(0.005)
    s.TimeStep(1)
    o.step(1)
    o32.step(1)
    for n in ("u1", "u2", "u3", "b"):
        want = o.get_modal(n)
        err, own = rel_l2(s.get(n), want), rel_l2(o32.get_modal(n), want)
        print("  %s GPU rel l2 %.3e, FP32 oracle %.3e" % (n, err, own), flush=True)
        assert err <= 3 * own + 1e-7, n



def test_full_size_transform_roundtrip_and_projection():
    import stratiflow_amd as sf
    N1, N2, N3, dim = C2
    s = sf.IMEXRK(N1, N2, N3, dim, 2 * np.pi / 0.44, np.pi / 0.44, 10.0, 4000.0, 0.16, 1.0, precision=4)
    rng = np.random.default_rng(5)
    shape = (N2, N1 // 2 + 1, N3)
    a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    a[0, 0, :] = a[0, 0, :].real  # the (0,0) stack of a real field
    for j2 in range(1, N2 // 2 + 1):  # Hermitian j1 = 0 column
        a[N2 - j2, 0, :] = np.conj(a[j2, 0, :])
    f = s.field(sf.MODAL, sf.NEUMANN)
    f.upload(a)
    ref = f.download()  # filtered
    g = s.field(sf.NODAL, sf.NEUMANN)
    f.ToNodal(g)
    h = s.field(sf.MODAL, sf.NEUMANN)
    g.ToModal(h)
    err = rel_l2(h.download(), ref)
    print("round trip rel l2 %.3e" % err)
    assert err < 2e-6

    # projection: div = ikx u1 + iky u2 + Dz u3 vanishes on interior rows
    from stratiflow_amd import ic
    z, zf = s.vertical_points()
    pert = ic.perturbation(N1, N2, N3, dim, z, zf, 3, amplitude=1e-2)
    for n, arr in pert.items():
        s.set(n, arr)
    s.FilterAll()
    s.PopulateNodalVariables()
    s.RemoveDivergence(0.0)
    d1, d2 = s.field(sf.MODAL, sf.NEUMANN), s.field(sf.MODAL, sf.NEUMANN)
    d3 = s.field(sf.MODAL, sf.NEUMANN)
    lib = sf.lib()
    s._chk(lib.sf_ddx(s.state("u1").h, d1.h))
    s._chk(lib.sf_ddy(s.state("u2").h, d2.h))
    s._chk(lib.sf_ddz(s.state("u3").h, d3.h))
    div = d1.download() + d2.download() + d3.download()
    scale = np.abs(d1.download()).max()
    worst = np.abs(div[..., 2:N3 - 2]).max() / scale
    print("max |div| / max |ddx u1| on interior rows %.3e" % worst)
    assert worst < 1e-4
