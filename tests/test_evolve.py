"""The facade's adaptive evolution loops against the oracle driven the same way.

tests/cpp/evolve_demo.cpp runs StateVector::FullEvolve (StateVector.cpp:3-83)
with snapshots and the mixing integral, an Adjoint.cpp-style pass over the
snapshots it wrote, and StateVector::LinearEvolve (StateVector.cpp:107-156) to
a time that is not a multiple of its step.  The oracle repeats both loops in
the reference's FP32 time bookkeeping (t, dt and the remaining-step count in
single precision), with its own CFL/CFLlinear, steps and JoverK in FP64.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle
from stratiflow_amd import ic

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "_build", "evolve_demo")
N1, N2, N3 = 48, 1, 64
L1 = 2 * 3.14159265358979 / 0.44
f32 = np.float32

pytestmark = pytest.mark.gpu


def _oracle():
    o = oracle.Oracle(N1, N2, N3, oracle.TWO_D, L1, 1.0, 10.0, 1000.0, 0.1, 1.0, precision=8)
    z, zf, _, _ = o.grid()
    o.set_background(ic.tanh_background(zf))
    return o


def _load(o, path):
    raw = np.fromfile(path, dtype=np.float32).astype(np.float64).reshape(4, N2, N1, N3)
    return {"u1": o.to_modal(raw[0]), "u3": o.to_modal(raw[2]), "b": o.to_modal(raw[3])}


def _start(o, st):
    for n, a in st.items():
        o.set_modal(n, a)
    o.set_modal("p", np.zeros(o.modal_shape, dtype=complex))
    # FullEvolve / LinearEvolve start-up, then PrepareRun's populate
    o.filter_all()
    o.populate_nodal()
    o.remove_divergence(0.0)
    o.populate_nodal()


def _norm(o):
    d = o.inner_prod(o.get_modal("u1"), o.get_modal("u1"), oracle.NEUMANN)
    d += o.inner_prod(o.get_modal("u3"), o.get_modal("u3"), oracle.DIRICHLET)
    d += 0.1 * o.inner_prod(o.get_modal("b"), o.get_modal("b"), oracle.NEUMANN)
    return np.sqrt(d)


def _evolve(o, T, linear, mixing=False):
    """The loop of StateVector.cpp:14-80 (linear: :116-153) with its FP32 times:
    CFL (CFLlinear) every 100 steps rescales dt to 0.8 (0.4), and when fewer
    than 100 steps remain dt is shortened so the last step ends at T."""
    T = f32(T)
    t, step, mix = f32(0), 0, f32(0)
    marks = []
    while t + f32(0.0001) < T:
        if step % 100 == 0:
            c = f32(o.cfl(linear=linear))
            dt = f32(f32(o.dt) * f32(0.4 if linear else 0.8) / c)
            o.set_dt(float(dt))
            remaining = f32(T - t)
            rs = int(remaining / dt) + 1
            if rs < 100:
                o.set_dt(float(f32(remaining / f32(rs))))
            marks.append((step, float(t)))
        if linear:
            o.step_linear(1)
        else:
            o.step(1)
        if mixing:
            mix = f32(mix + f32(o.jover_k()) * f32(o.dt))
        t = f32(t + f32(o.dt))
        step += 1
    marks.append((step, float(t)))
    return mix, marks


def test_full_and_linear_evolve_match_oracle(tmp_path):
    out = str(tmp_path / "ev")
    r = subprocess.run([BIN, out, "60", "0.537"], capture_output=True, text=True, timeout=300, cwd=str(tmp_path))
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    vals = {}
    for line in r.stdout.strip().splitlines():
        p = line.split()
        if len(p) == 2:
            vals[p[0]] = float(p[1])
    print({k: v for k, v in vals.items() if not k.startswith("snap_")}, flush=True)

    o = _oracle()
    init = _load(o, out + "_init.fields")
    _start(o, init)
    mix, snaps = _evolve(o, 60.0, False, mixing=True)
    ke, pe = o.energy()
    print("oracle: mixing %.9e ke %.9e pe %.9e dt %.9e snaps %s" % (mix, ke, pe, o.dt, snaps), flush=True)
    # same snapshot steps and times (StateVector.cpp:46-53, 68-76)
    assert int(vals["snapshots"]) == len(snaps)
    for step, t in snaps:
        assert "snap_%d" % step in vals, step
        assert abs(vals["snap_%d" % step] - t) < 1e-4 * max(1.0, t)
    assert abs(vals["final_ke"] - ke) / ke < 1e-4
    assert abs(vals["final_pe"] - pe) / pe < 1e-4
    assert abs(vals["final_dt"] - o.dt) / o.dt < 1e-5
    assert abs(vals["mixing"] - float(mix)) / abs(float(mix)) < 1e-4
    assert vals["last_snapshot_vs_result"] < 1e-6
    assert vals["adjoint_interval_steps"] >= 1 and np.isfinite(vals["adjoint_norm"])

    # LinearEvolve from the initial state about the final state
    about = {n: o.get_modal(n) for n in ("u1", "u3", "b")}
    lo = _oracle()
    for n in ("u1", "u3", "b"):
        lo.set_modal(n + "t", about[n])
    lo.set_linear_background()
    _start(lo, init)
    lo.set_dt(0.01)
    _, lmarks = _evolve(lo, 0.537, True)
    lnorm = _norm(lo)
    print("oracle linear: norm %.9e dt %.9e steps %d" % (lnorm, lo.dt, lmarks[-1][0]), flush=True)
    # ends exactly at T: the last steps are shortened
    assert abs(lmarks[-1][1] - 0.537) < 1e-5
    assert abs(vals["linear_dt"] - lo.dt) / lo.dt < 1e-5
    assert abs(vals["linear_norm"] - lnorm) / lnorm < 1e-4
