"""The batched ensemble workload: 64 independent 2D runs in one batch.

N1 drawn from {64, 96, 128, 192, 256} and N3 from {96, 128, 192, 256, 384},
Re uniform in [500, 4000], Ri uniform in [0.05, 0.25], Pr = 1, each run with its
own seeded perturbation of amplitude 1e-3, dt = 0.01 (the generator bench.py's
--ensemble uses).  A bounded number of steps: every batched run must equal its
twin context stepped on its own bit for bit (to 1e-4 of the field's largest
value if the tuner picked the x-pass FFT plans with first radix 8, which round
differently), and the runs on the largest grid drawn must match the FP32 and
FP64 oracles' energies to 1e-4.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SEED = 20261202
STEPS = 40


def specs(nruns=64, seed=SEED):
    rng = np.random.default_rng(seed)
    return [(int(rng.choice([64, 96, 128, 192, 256])), int(rng.choice([96, 128, 192, 256, 384])),
             float(rng.uniform(500, 4000)), float(rng.uniform(0.05, 0.25))) for _ in range(nruns)]


def _make(k, N1, N3, Re, Ri):
    import stratiflow_amd as sf
    from stratiflow_amd import ic
    L1 = 2 * np.pi / 0.44
    s = sf.IMEXRK(N1, 1, N3, sf.TWO_D, L1, 1.0, 10.0, Re, Ri, 1.0, precision=4)
    z, zf = s.vertical_points()
    s.SetBackground(ic.tanh_background(zf))
    for n, a in ic.perturbation(N1, 1, N3, sf.TWO_D, z, zf, SEED + k, amplitude=1e-3).items():
        s.set(n, a)
    s.FilterAll()
    s.PopulateNodalVariables()
    s.RemoveDivergence(0.0)
    s.deltaT = 0.01
    return s


def test_ensemble_64_runs_bitwise_and_oracle():
    import oracle
    import stratiflow_amd as sf
    from stratiflow_amd import ic
    sp = specs()
    grids = sorted(set((a, b) for a, b, _, _ in sp))
    print("%d runs on %d distinct grids" % (len(sp), len(grids)), flush=True)
    runs = [_make(k, *s) for k, s in enumerate(sp)]
    twins = [_make(k, *s) for k, s in enumerate(sp)]
    b = sf.Batch(runs)
    info = b.info()
    print("batch", info, flush=True)
    exact = info["xpass_plan"] != "mixed_r8"  # another FFT plan rounds differently
    b.step(STEPS)
    for t in twins:
        t.TimeStep(STEPS)
    for k, (s, t) in enumerate(zip(runs, twins)):
        for f in ("u1", "u3", "b", "p"):
            a, c = s.get(f), t.get(f)
            assert np.isfinite(a).all(), (k, f)
            if exact:
                assert np.array_equal(a, c), (k, sp[k], f)
            else:
                assert np.abs(a - c).max() <= 1e-4 * np.abs(c).max(), (k, sp[k], f)
    # the largest grid drawn against the oracles
    big = max(range(len(sp)), key=lambda k: sp[k][0] * sp[k][1])
    N1, N3, Re, Ri = sp[big]
    ke, pe = runs[big].energies()
    L1 = 2 * np.pi / 0.44
    for prec in (4, 8):
        o = oracle.Oracle(N1, 1, N3, oracle.TWO_D, L1, 1.0, 10.0, Re, Ri, 1.0, precision=prec)
        z, zf, _, _ = o.grid()
        o.set_background(ic.tanh_background(zf))
        for n, a in ic.perturbation(N1, 1, N3, oracle.TWO_D, z, zf, SEED + big, amplitude=1e-3).items():
            o.set_modal(n, a)
        o.filter_all()
        o.populate_nodal()
        o.remove_divergence(0.0)
        o.set_dt(0.01)
        o.step(STEPS)
        ko, po = o.energy()
        print("run %d %s: KE gpu %.9e oracle(fp%d) %.9e  PE gpu %.9e oracle %.9e"
              % (big, sp[big], ke, 8 * prec, ko, pe, po), flush=True)
        assert abs(ke - ko) / ko < 1e-4 and abs(pe - po) / po < 1e-4
    b.close()
    for r in runs + twins:
        r.close()
