"""Golden fixtures (tests/golden/*.npz, written by make_golden.py from the FP64 oracle).

CPU: the oracle reproduces them (regression guard).
GPU: the HIP path reproduces them through the C ABI, FP64 tightly and FP32 to
the 1e-4 relative energy bar.
"""
import ast
import glob
import os

import numpy as np
import pytest

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
# the state fixtures of make_golden.py (c2d_*, c3d_*), not the energy series
FILES = sorted(glob.glob(os.path.join(HERE, "golden", "c[23]d_*.npz")))


def load(path):
    d = np.load(path, allow_pickle=False)
    meta = ast.literal_eval(str(d["meta"]))
    return meta, d


def names(meta):
    return ["u1", "u2", "u3", "b", "p"] if meta["dim"] == 0 else ["u1", "u3", "b", "p"]


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


@pytest.mark.parametrize("path", FILES, ids=[os.path.basename(f) for f in FILES])
def test_oracle_reproduces_fixture(path):
    meta, d = load(path)
    o = oracle.Oracle(meta["N1"], meta["N2"], meta["N3"], meta["dim"], meta["L1"], meta["L2"], meta["L3"],
                      meta["Re"], meta["Ri"], meta["Pr"], precision=8)
    o.set_background(d["U"], d["S"])
    for n in ["u1", "u2", "u3", "b"]:
        if "in_" + n in d:
            o.set_modal(n, d["in_" + n])
    o.filter_all()
    o.populate_nodal()
    o.remove_divergence(0.0)
    o.set_dt(meta["dt"])
    o.step(10)
    for n in names(meta):
        assert rel(o.get_modal(n), d["step10_" + n]) < 1e-12, n


def _gpu_run(meta, d, prec, steps):
    import stratiflow_amd as sf
    s = sf.IMEXRK(meta["N1"], meta["N2"], meta["N3"], meta["dim"], meta["L1"], meta["L2"], meta["L3"],
                  meta["Re"], meta["Ri"], meta["Pr"], precision=prec)
    s.SetBackground(d["U"], d["S"])
    for n in ["u1", "u2", "u3", "b"]:
        if "in_" + n in d:
            s.set(n, d["in_" + n])
    s.FilterAll()
    s.PopulateNodalVariables()
    s.RemoveDivergence(0.0)
    s.deltaT = meta["dt"]
    return s


@pytest.mark.gpu
@pytest.mark.parametrize("path", FILES, ids=[os.path.basename(f) for f in FILES])
def test_gpu_fp64_matches_fixture(path):
    meta, d = load(path)
    s = _gpu_run(meta, d, 8, 0)
    for n in names(meta):
        assert rel(s.get(n), d["proj_" + n]) < 1e-11, ("proj", n)
    s.Stage(0)
    for n in names(meta) + ["r1", "r3", "rB"]:
        assert rel(s.get(n), d["stage0_" + n]) < 1e-10, ("stage0", n)
    s.Stage(1)
    s.Stage(2)
    for n in names(meta):
        assert rel(s.get(n), d["step1_" + n]) < 1e-10, ("step1", n)
    s.TimeStep(9)
    for n in names(meta):
        assert rel(s.get(n), d["step10_" + n]) < 1e-9, ("step10", n)
    ke, pe = s.energies()
    assert abs(ke - d["step10_energy"][0]) / d["step10_energy"][0] < 1e-10
    assert abs(pe - d["step10_energy"][1]) / d["step10_energy"][1] < 1e-10
    s.PopulateNodalVariables()
    assert rel(s.get("U1"), d["nodal_U1_step10"]) < 1e-10


@pytest.mark.gpu
@pytest.mark.parametrize("path", FILES, ids=[os.path.basename(f) for f in FILES])
def test_gpu_fp32_matches_fixture(path):
    meta, d = load(path)
    s = _gpu_run(meta, d, 4, 0)
    s.TimeStep(10)
    ke, pe = s.energies()
    eke = abs(ke - d["step10_energy"][0]) / d["step10_energy"][0]
    epe = abs(pe - d["step10_energy"][1]) / d["step10_energy"][1]
    print("fp32 vs fp64 fixture: KE %.2e PE %.2e" % (eke, epe))
    assert eke < 1e-4 and epe < 1e-4
    for n in names(meta):
        assert rel(s.get(n), d["step10_" + n]) < 1e-4, n
