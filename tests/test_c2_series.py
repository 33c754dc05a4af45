"""The N-step energy criterion on the throughput workload.

384x192x384 FP32, Re=4000, Ri=0.16, Pr=1, U = tanh z, B = tanh z, the seeded
perturbation of amplitude 1e-3 (seed 20261202), dt = 0.005 -- the grid, flow
and initial state bench.py times.  FullEvolve's start-up (FilterAll,
PopulateNodalVariables, RemoveDivergence(0), StateVector.cpp:9-21) and 200
TimeSteps; KE and PE (IMEXRK.h:395-411) after every step must stay within 1e-4
relative of the FP32 oracle's series (tests/golden/c2_series_384x192x384.npz,
written by make_c2_series.py), and the error against the FP64 oracle's series
is printed beside it.
"""
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "c2_series_384x192x384.npz")
N1, N2, N3, STEPS = 384, 192, 384, 200


def _fixture():
    if not os.path.exists(FIXTURE):
        pytest.skip("c2 series fixture not generated")
    return np.load(FIXTURE)


def test_fixture_series_fp32_against_fp64():
    """The FP32 oracle's series follows the FP64 one (the FP32 restatement of
    the path is stable over the 200 steps) and the energy decays smoothly."""
    d = _fixture()
    for k in ("ke32", "pe32", "ke64", "pe64"):
        assert d[k].shape == (STEPS + 1,) and np.all(d[k] > 0), k
    for a, b in (("ke32", "ke64"), ("pe32", "pe64")):
        e = np.abs(d[a] - d[b]) / d[b]
        print("%s vs %s: max rel %.3e" % (a, b, e.max()))
        assert e.max() < 1e-4
    assert d["ke64"][-1] < d["ke64"][0]


@pytest.mark.gpu
def test_c2_energy_series_200_steps_against_oracle():
    import stratiflow_amd as sf
    from stratiflow_amd import ic
    d = _fixture()
    L1 = 2 * np.pi / 0.44
    s = sf.IMEXRK(N1, N2, N3, sf.THREE_D, L1, L1 / 2, 10.0, 4000.0, 0.16, 1.0, precision=4)
    z, zf = s.vertical_points()
    s.SetBackground(ic.tanh_background(zf), ic.tanh_stratification(zf))
    for n, a in ic.perturbation(N1, N2, N3, sf.THREE_D, z, zf, 20261202, amplitude=1e-3).items():
        s.set(n, a)
    s.FilterAll()
    s.PopulateNodalVariables()
    s.RemoveDivergence(0.0)
    s.deltaT = 0.005
    ke, pe = [], []
    k, p = s.energies()
    ke.append(k)
    pe.append(p)
    for step in range(STEPS):
        s.TimeStep(1)
        k, p = s.energies()
        ke.append(k)
        pe.append(p)
        if step % 50 == 49:
            print("step %d KE %.9e (oracle %.9e) PE %.9e (oracle %.9e)" % (step + 1, k, d["ke32"][step + 1], p,
                                                                       d["pe32"][step + 1]), flush=True)
    s.close()
    ke, pe = np.array(ke), np.array(pe)
    for name, g in (("KE", ke), ("PE", pe)):
        e32 = np.abs(g - d[name.lower() + "32"]) / d[name.lower() + "32"]
        e64 = np.abs(g - d[name.lower() + "64"]) / d[name.lower() + "64"]
        print("%s over 200 steps: max rel vs FP32 oracle %.3e (at step %d), vs FP64 oracle %.3e; final %.3e / %.3e"
              % (name, e32.max(), int(e32.argmax()), e64.max(), e32[-1], e64[-1]), flush=True)
        assert e32.max() < 1e-4, name
        assert e64.max() < 1e-4, name
