"""Batched ensemble (sf_batch_*): many independent 2D runs advanced together.

Each batched run must end bit-identical to a twin context stepped on its own
(TimeStep, IMEXRK.cpp:8-24): the batched kernels run the same per-level and
per-mode code, only the assignment of work to blocks differs.  The exception
is an x-pass FFT plan other than the runs' own (first radix 8): it rounds
differently, as FFTW's timed plans do (FFTW_PATIENT, Field.h:730-991), so those
runs agree with their twins to rounding (FP32 1e-5, FP64 1e-12 of the field's
largest value after a few steps, ten times that for the pressure, or in FP32
three times the distance of the FP32 twin from an FP64 twin, whichever is
larger: the pressure of a small run is fixed only to a few 1e-4 of its maximum
by FP32 rounding).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# (N1, N3, Re, Ri, dt): uneven grids and parameters, two x-pass groups share N1
RUNS = [(32, 48, 500.0, 0.05, 0.01), (48, 64, 1200.0, 0.12, 0.01), (32, 96, 2000.0, 0.2, 0.005),
        (64, 128, 4000.0, 0.25, 0.01), (48, 160, 800.0, 0.1, 0.008), (96, 64, 3000.0, 0.15, 0.01)]


def same(a, c, exact, prec, field="u1", c64=None):
    """Bit-identical, or within rounding of the other FFT plan (ten times
    looser for the pressure, which the projection rescales by 1/h; in FP32 at
    least three times the twin's own distance from its FP64 twin c64)."""
    if exact:
        return np.array_equal(a, c)
    tol = (1e-5 if prec == 4 else 1e-12) * (10 if field == "p" else 1) * np.abs(c).max()
    if c64 is not None:
        tol = max(tol, 3 * np.abs(c - c64).max())
    return np.abs(a - c).max() <= tol


def _make(N1, N3, Re, Ri, dt, seed, prec):
    import stratiflow_amd as sf
    from stratiflow_amd import ic
    L1 = 2 * np.pi / 0.44
    s = sf.IMEXRK(N1, 1, N3, sf.TWO_D, L1, 1.0, 10.0, Re, Ri, 1.0, precision=prec)
    z, zf = s.vertical_points()
    s.SetBackground(ic.tanh_background(zf))
    for n, a in ic.perturbation(N1, 1, N3, sf.TWO_D, z, zf, seed, amplitude=1e-2).items():
        s.set(n, a)
    s.FilterAll()
    s.PopulateNodalVariables()
    s.RemoveDivergence(0.0)
    s.deltaT = dt
    return s


@pytest.mark.parametrize("prec", [4, 8])
def test_batch_matches_individual_runs(prec):
    import stratiflow_amd as sf
    batched = [_make(*r, seed=100 + i, prec=prec) for i, r in enumerate(RUNS)]
    twins = [_make(*r, seed=100 + i, prec=prec) for i, r in enumerate(RUNS)]
    b = sf.Batch(batched)
    info = b.info()
    print("batch", info, flush=True)
    assert info["xpass_groups"] == 4 and info["modal_groups"] == 3
    exact = info["xpass_plan"] != "mixed_r8"
    # FP64 twins measure the FP32 rounding that another FFT plan may change
    twins64 = [_make(*r, seed=100 + i, prec=8) for i, r in enumerate(RUNS)] if (prec == 4 and not exact) else []
    b.step(5)
    for t in twins + twins64:
        t.TimeStep(5)
    b.step(3)  # graph replays only
    for t in twins + twins64:
        t.TimeStep(3)
    for i, (s, t) in enumerate(zip(batched, twins)):
        for f in ("u1", "u3", "b", "p"):
            a, c = s.get(f), t.get(f)
            c64 = twins64[i].get(f) if twins64 else None
            assert np.isfinite(a).all()
            assert same(a, c, exact, prec, f, c64), (i, f, np.abs(a - c).max() / np.abs(c).max())
    # a dt change is picked up (the graph is rebuilt)
    batched[0].deltaT = 0.004
    twins[0].deltaT = 0.004
    if twins64:
        twins64[0].deltaT = 0.004
    b.step(2)
    for t in twins + twins64:
        t.TimeStep(2)
    for i, (s, t) in enumerate(zip(batched, twins)):
        assert same(s.get("u1"), t.get("u1"), exact, prec, "u1", twins64[i].get("u1") if twins64 else None)
    b.close()


@pytest.mark.parametrize("prec", [4, 8])
def test_every_batch_plan_matches_individual_runs(prec):
    """One launch per N1 / per vertical tile, or one launch for all runs (the
    x-pass with the default FFT plans or first radix 8, the modal kernel with 4
    or 8 modes per block): the tuner's choice changes the time only."""
    import stratiflow_amd as sf
    batched = [_make(*r, seed=200 + i, prec=prec) for i, r in enumerate(RUNS)]
    twins = [_make(*r, seed=200 + i, prec=prec) for i, r in enumerate(RUNS)]
    twins64 = [_make(*r, seed=200 + i, prec=8) for i, r in enumerate(RUNS)] if prec == 4 else []
    b = sf.Batch(batched)
    info = b.info()
    print("tuned", info, flush=True)
    t = info["tuned_ms_per_stage"]
    assert all(v > 0 for v in t.values())
    xbest = min(("grouped", "mixed", "mixed_r8"), key=lambda k: t["xpass_" + k])
    mbest = min(("grouped", "mixed4", "mixed8"), key=lambda k: t["modal_" + k])
    assert (info["xpass_plan"], info["modal_plan"]) == (xbest, mbest)
    # the plans with the runs' own FFT plans first (bit for bit), then first radix 8
    plans = [(sf.BATCH_MIXED, sf.BATCH_MIXED), (sf.BATCH_GROUPED, sf.BATCH_MIXED8), (sf.BATCH_GROUPED, sf.BATCH_GROUPED),
             (sf.BATCH_MIXED, sf.BATCH_MIXED8), (sf.BATCH_MIXED_R8, sf.BATCH_GROUPED), (sf.BATCH_MIXED_R8, sf.BATCH_MIXED)]
    exact = True
    for x, m in plans:
        b.select(x, m)
        info = b.info()
        assert info["launches_per_step"] == 3 * ((1 if x else 4) + (1 if m else 3)), info
        exact = exact and x != sf.BATCH_MIXED_R8
        b.step(2)
        for tw in twins + twins64:
            tw.TimeStep(2)
        for i, (s, tw) in enumerate(zip(batched, twins)):
            for f in ("u1", "u3", "b", "p"):
                a, c = s.get(f), tw.get(f)
                c64 = twins64[i].get(f) if twins64 else None
                assert np.isfinite(a).all()
                assert same(a, c, exact, prec, f, c64), (x, m, i, f, np.abs(a - c).max() / np.abs(c).max())
    b.close()


def test_radix8_plans_of_larger_grids_fp64():
    """The first-radix-8 x FFT plans of N1 = 128, 192 and 256 (8 x 16, 8 x 24,
    8 x 32: lane pairs of radix 8, 12 and 16) against the runs' own default
    plans (16 x 8, 16 x 12, 16 x 16), in FP64 after four steps."""
    import stratiflow_amd as sf
    runs = [(128, 64, 1500.0, 0.1, 0.01), (192, 96, 2500.0, 0.15, 0.01), (256, 48, 800.0, 0.2, 0.008)]
    batched = [_make(*r, seed=300 + i, prec=8) for i, r in enumerate(runs)]
    twins = [_make(*r, seed=300 + i, prec=8) for i, r in enumerate(runs)]
    b = sf.Batch(batched)
    b.select(sf.BATCH_MIXED_R8, sf.BATCH_MIXED)
    assert b.info()["xpass_plan"] == "mixed_r8"
    b.step(4)
    for t in twins:
        t.TimeStep(4)
    for i, (s, t) in enumerate(zip(batched, twins)):
        for f in ("u1", "u3", "b", "p"):
            a, c = s.get(f), t.get(f)
            err = np.abs(a - c).max() / np.abs(c).max()
            print(runs[i][:2], f, "%.2e" % err, flush=True)
            assert err <= (1e-11 if f == "p" else 1e-12), (i, f, err)
    b.close()
