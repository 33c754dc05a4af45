#!/usr/bin/env python3
"""Throughput of the IMEX-RK3/CN timestep (IMEXRK::TimeStep, IMEXRK.cpp:8-24) on one MI355X.

Workload: 3D stratified shear layer, 384 (x, Fourier) x 192 (y, Fourier) x 384
(z, stretched finite differences), FP32, Lx = 2*pi/0.44, Ly = Lx/2, L3 = 10,
Re = 4000, Pr = 1, Ri = 0.16, U = tanh(z), B = tanh(z), seeded perturbation of
amplitude 1e-3, dt = 0.005.  One "step" is one full three-stage TimeStep.

Prints one JSON line.  ``value`` is steps/s with the state resident in HBM,
timed with HIP events on the solver's stream.  ``roofline`` is the dominant
kernel's algorithmic bytes per launch over its average launch time;
``step_roofline`` applies the per-step byte model to the whole step;
``cpu_baseline`` times the CPU oracle (a C++ restatement of the reference,
OpenMP) on the host cores for a bounded sample of the same workload.
``newton_gmres`` (after the timed region; tests/cpp/newton_bench.cpp, compiled
here when its binary is missing) runs Newton-GMRES toward relative residual
1e-8 on held-out 256x1x256 problems with plain GMRES, the analytic
preconditioner and the learned network on it: Newton steps, operator
applications, wall time to each residual decade, final residuals.
"""
import subprocess
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0
METRIC = "steps/s @384×192×384 fp32; Newton-GMRES time, learned vs plain @1e-8 res"
GRID = dict(N1=384, N2=192, N3=384)
FLOW = dict(Re=4000.0, Ri=0.16, Pr=1.0, L3=10.0)
DT = 0.005
SEED = 20261202


def sizes(N1, N2, N3, word=4):
    A1 = N1 // 3
    A2 = 2 * (N2 // 3) - 1 if N2 > 1 else 1
    cplx = 2 * word
    Ba = A1 * A2 * N3 * cplx          # one active modal field
    Bi = N2 * A1 * N3 * cplx          # one x-spectral field (all n2, active j1)
    Bn = N1 * N2 * N3 * word          # one nodal field
    return A1, A2, Ba, Bi, Bn


def kernel_bytes(N1, N2, N3, word=4):
    """Algorithmic (compulsory) HBM bytes per launch of each kernel class, 3D."""
    _, _, Ba, Bi, _ = sizes(N1, N2, N3, word)
    # modal: per stage reads nl(4)+u(4)+p, r_old(4) on stages 1,2; writes u(4)+p, r_new(4) on stages 0,1
    modal_avg = ((9 + 13 + 13) + (9 + 9 + 5)) / 3.0 * Ba
    return {"yinv": 4 * Ba + 4 * Bi, "xpass": 4 * Bi + 8 * Bi, "yfwd": 8 * Bi + 4 * Ba, "modal": modal_avg}


def model_step_bytes(N1, N2, N3, word=4):
    """SF-IO-v1 byte model of one step: 3*[Bn(2nF+2nP) + Ba(5nF+2nP+2)]."""
    _, _, Ba, _, Bn = sizes(N1, N2, N3, word)
    nF, nP = (4, 11) if N2 > 1 else (3, 6)
    return 3 * (Bn * (2 * nF + 2 * nP) + Ba * (5 * nF + 2 * nP + 2))


def setup(sf, ic, prec=4):
    import numpy as np
    N1, N2, N3 = GRID["N1"], GRID["N2"], GRID["N3"]
    L1 = 2 * np.pi / 0.44
    s = sf.IMEXRK(N1, N2, N3, sf.THREE_D, L1, L1 / 2, FLOW["L3"], FLOW["Re"], FLOW["Ri"], FLOW["Pr"], precision=prec)
    z, zf = s.vertical_points()
    U, S = ic.tanh_background(zf), ic.tanh_stratification(zf)
    pert = ic.perturbation(N1, N2, N3, sf.THREE_D, z, zf, SEED, amplitude=1e-3)
    return s, U, S, pert


def cpu_baseline(U, S, pert, max_seconds):
    """Time the CPU oracle (FP32, OpenMP) on one step of the same workload."""
    import numpy as np
    import oracle
    ncores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    ncores = max(1, min(ncores, 16))
    oracle.set_threads(ncores)
    N1, N2, N3 = GRID["N1"], GRID["N2"], GRID["N3"]
    L1 = 2 * np.pi / 0.44
    o = oracle.Oracle(N1, N2, N3, oracle.THREE_D, L1, L1 / 2, FLOW["L3"], FLOW["Re"], FLOW["Ri"], FLOW["Pr"],
                      precision=4)
    o.set_background(U, S)
    for n, a in pert.items():
        o.set_modal(n, a)
    o.filter_all()
    o.populate_nodal()
    o.remove_divergence(0.0)
    o.set_dt(DT)
    t0 = time.perf_counter()
    o.stage(0)
    t_stage = time.perf_counter() - t0
    nstages = 1
    # complete the step if it fits the budget
    if 2 * t_stage < max_seconds:
        t0 = time.perf_counter()
        o.stage(1)
        o.stage(2)
        t_stage = (t_stage + time.perf_counter() - t0) / 3.0
        nstages = 3
    step_s = 3.0 * t_stage
    cpu = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"value": 1.0 / step_s, "unit": "steps/s", "cores": ncores, "kind": "port",
            "sample": "%d RK stage(s) of one 384x192x384 FP32 step of the CPU oracle (C++14/OpenMP restatement, own "
                      "FFT, banded vertical operators), setup excluded; %s" % (nstages, cpu)}


def ensemble(sf, ic, nruns=64, steps=500, seed=20261202):
    """Batched 2D ensemble: nruns seeded runs with uneven grids and parameters,
    advanced together (sf_batch_*) and, for comparison, one after the other."""
    import numpy as np
    rng = np.random.default_rng(seed)
    specs = [(int(rng.choice([64, 96, 128, 192, 256])), int(rng.choice([96, 128, 192, 256, 384])),
              float(rng.uniform(500, 4000)), float(rng.uniform(0.05, 0.25))) for _ in range(nruns)]

    def make(k, N1, N3, Re, Ri):
        L1 = 2 * np.pi / 0.44
        s = sf.IMEXRK(N1, 1, N3, sf.TWO_D, L1, 1.0, 10.0, Re, Ri, 1.0, precision=4)
        z, zf = s.vertical_points()
        s.SetBackground(ic.tanh_background(zf))
        for n, a in ic.perturbation(N1, 1, N3, sf.TWO_D, z, zf, seed + k, amplitude=1e-3).items():
            s.set(n, a)
        s.FilterAll()
        s.PopulateNodalVariables()
        s.RemoveDivergence(0.0)
        s.deltaT = 0.01
        return s

    runs = [make(k, *sp) for k, sp in enumerate(specs)]
    b = sf.Batch(runs)
    b.step(1)  # first step (nodal snapshots) and graph capture
    ms_batch = b.time_steps(steps)
    seq = [make(k, *sp) for k, sp in enumerate(specs)]
    for r in seq:
        r.TimeStep(1)
        r.synchronize()
    t0 = time.perf_counter()
    for r in seq:
        r.TimeStep(steps)
    for r in seq:
        r.synchronize()
    ms_seq = (time.perf_counter() - t0) * 1e3
    info = b.info()
    out = {"runs": nruns, "steps": steps, "grids": "N1 in {64,96,128,192,256}, N3 in {96,128,192,256,384}",
           "batched_run_steps_per_s": nruns * steps / (ms_batch * 1e-3),
           "sequential_run_steps_per_s": nruns * steps / (ms_seq * 1e-3),
           "speedup": ms_seq / ms_batch, "launches_per_step": info["launches_per_step"],
           "xpass_groups": info["xpass_groups"], "modal_groups": info["modal_groups"],
           "xpass_plan": info["xpass_plan"], "modal_plan": info["modal_plan"],
           "tuned_ms_per_stage": {k: round(v, 5) for k, v in info["tuned_ms_per_stage"].items()}}
    b.close()
    return out


NEWTON_ARGS = ["ntrain=2", "ntest=2", "newton=10", "batches=3000"]


def newton_binary():
    """tests/cpp/newton_bench.cpp built against the in-tree library (FP64
    stepping); built here when missing.  Returns (path, error)."""
    exe = os.path.join(HERE, "tests", "cpp", "_build", "newton_bench_f64")
    if os.path.exists(exe):
        return exe, None
    os.makedirs(os.path.dirname(exe), exist_ok=True)
    lib = os.path.join(HERE, "stratiflow_amd")
    try:
        r = subprocess.run(["g++", "-std=c++14", "-O2", "-DUSE_DOUBLE", "-I" + os.path.join(HERE, "include"), "-o",
                            exe, os.path.join(HERE, "tests", "cpp", "newton_bench.cpp"), "-L" + lib, "-lsfhip",
                            "-Wl,-rpath," + lib], capture_output=True, text=True, timeout=300)
    except (OSError, subprocess.TimeoutExpired) as e:
        return None, "building newton_bench failed: %s" % e
    if r.returncode != 0:
        return None, "building newton_bench failed: " + r.stderr[-400:]
    return exe, None


def newton_gmres(timeout=480):
    """Newton-GMRES toward |F| <= 1e-8 |F(x0)| on held-out problems of the
    learned-preconditioner class (256x1x256, Re in [1000, 3000], Ri in
    [0.08, 0.2], F(x) = Phi_2(x) - x, FP64 stepping, Krylov dimension 60, GMRES
    tolerance 1e-2, at most 10 Newton steps): plain GMRES, the analytic
    advected-diffusion preconditioner, and the learned network on that base
    (trained on the GPU on the operator applications of two plain training
    runs), judged by plain GMRES."""
    exe, err = newton_binary()
    if err:
        return {"error": err}
    try:
        r = subprocess.run([exe] + NEWTON_ARGS, capture_output=True, text=True, timeout=timeout)
    except subprocess.TimeoutExpired:
        return {"error": "newton_bench exceeded %d s" % timeout}
    except OSError as e:
        return {"error": "newton_bench could not run: %s" % e}
    if r.returncode != 0:
        return {"error": (r.stdout + r.stderr)[-400:]}
    out = summarize_newton(r.stdout)
    out["workload"] = ("Newton-GMRES to relative residual 1e-8 (at most 10 Newton steps), 2 held-out problems, "
                       "256x1x256 2D, Re~U[1000,3000], Ri~U[0.08,0.2], F(x)=Phi_2(x)-x, FP64 stepping, K=60, "
                       "GMRES tol 1e-2; learned = 4 x Conv1D(64 ch, k=5, dilations 1,2,4,8) BF16/MFMA network "
                       "correcting the analytic advected-diffusion preconditioner (delta 0.3), trained on the "
                       "GPU on the operator applications of 2 plain training runs")
    return out


def summarize_newton(text):
    """Totals per method over the held-out problems of newton_bench's output
    (its test/hist/loss lines): Newton steps, operator applications, seconds,
    final residuals, fallbacks, the summed time to each residual decade (None
    when some problem never reaches it) and how many problems reach it."""
    methods = {}
    loss = None
    for line in text.splitlines():
        t = line.split()
        if not t:
            continue
        if t[0] == "loss":
            loss = {"first": float(t[1]), "last": float(t[2]), "samples": int(t[3]), "train_s": float(t[4])}
        elif t[0] == "test":
            m = methods.setdefault(t[4], {"newton_steps": 0, "matvecs": 0, "seconds": 0.0, "final_rel_residual": [],
                                          "fallbacks": 0, "time_to": {}, "problems_reaching": {}})
            m["newton_steps"] += int(t[5])
            m["matvecs"] += int(t[6])
            m["seconds"] += float(t[7])
            m["final_rel_residual"].append(float(t[8]))
            m["fallbacks"] += int(t[9])
        elif t[0] == "hist":
            m = methods[t[2]]
            for e in range(2, 9):
                thr = 10.0 ** -e
                hit = next((float(p.split("@")[1]) for p in t[3:] if float(p.split("@")[0]) <= thr), None)
                key = "1e-%d" % e
                m["problems_reaching"][key] = m["problems_reaching"].get(key, 0) + (hit is not None)
                if hit is None:
                    m["time_to"][key] = None
                elif m["time_to"].get(key, 0.0) is not None:
                    m["time_to"][key] = m["time_to"].get(key, 0.0) + hit
    for m in methods.values():
        m["reached_1e-8"] = m["time_to"].get("1e-8") is not None
    # the smallest residual decade that plain and learned both reach on every problem
    ratio, level = None, None
    p, l = methods.get("plain"), methods.get("learned")
    if p and l:
        for e in range(8, 1, -1):
            key = "1e-%d" % e
            if p["time_to"].get(key) and l["time_to"].get(key):
                ratio, level = p["time_to"][key] / l["time_to"][key], key
                break
    return {
        "methods": methods,
        "training_loss": loss,
        "time_ratio_plain_over_learned": ratio,
        "at_residual": level,
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-newton", action="store_true", help="skip the Newton-GMRES comparison")
    ap.add_argument("--cpu-seconds", type=float, default=40.0, help="budget for the CPU baseline sample")
    ap.add_argument("--traffic-json", default=os.path.join(HERE, "profiles", "traffic_384x192x384_fp32.json"))
    ap.add_argument("--ensemble", action="store_true",
                    help="also time the batched 64-run 2D ensemble against sequential runs")
    ap.add_argument("--calibration-copies", type=int, default=0,
                    help="device-to-device copies of one modal field (known bytes) after the timed region, "
                         "to scale FETCH_SIZE/WRITE_SIZE counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the IMEX-RK step runs on a single GPU and is not sharded "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import stratiflow_amd as sf
    from stratiflow_amd import ic

    s, U, S, pert = setup(sf, ic)
    s.SetBackground(U, S)
    for n, a in pert.items():
        s.set(n, a)
    s.FilterAll()
    s.PopulateNodalVariables()
    s.RemoveDivergence(0.0)
    s.deltaT = DT
    s.TimeStep(args.warmup)
    s.synchronize()

    t0 = time.perf_counter()
    total_ms, per = s.time_steps_detailed(args.steps)
    s.synchronize()
    wall_ms = (time.perf_counter() - t0) * 1e3
    ms_step = total_ms / args.steps
    if args.calibration_copies:
        a = s.field(sf.MODAL, sf.NEUMANN)
        b = s.field(sf.MODAL, sf.NEUMANN)
        for _ in range(args.calibration_copies):
            b.assign(a)
        s.synchronize()
        _, _, Ba, _, _ = sizes(GRID["N1"], GRID["N2"], GRID["N3"])
        print("calibration: %d copies of %d bytes" % (args.calibration_copies, Ba), file=sys.stderr)
    ke, pe = s.energies()
    if not (np.isfinite(ke) and np.isfinite(pe)):
        raise RuntimeError("non-finite energy after the timed steps")

    N1, N2, N3 = GRID["N1"], GRID["N2"], GRID["N3"]
    kb = kernel_bytes(N1, N2, N3)
    dom = max(per, key=lambda k: per[k][0])
    dom_ms, dom_n = per[dom]
    achieved = kb[dom] / (dom_ms / dom_n * 1e-3) / 1e9
    traffic = None
    if os.path.exists(args.traffic_json):
        try:
            traffic = json.load(open(args.traffic_json)).get(dom)
        except (OSError, ValueError):
            traffic = None
    step_bytes = model_step_bytes(N1, N2, N3)
    step_gbs = step_bytes / (ms_step * 1e-3) / 1e9

    out = {
        "metric": METRIC,
        "value": 1e3 / ms_step,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "IMEX-RK3/CN TimeStep, 3D stratified shear layer 384x192x384 (x,y Fourier; z "
                               "stretched FD), Re=4000 Ri=0.16 Pr=1, tanh U and B, dt=0.005, seeded 1e-3 "
                               "perturbation", "N1": N1, "N2": N2, "N3": N3, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "bytes_per_launch": kb[dom], "ms_per_launch": dom_ms / dom_n},
        "step_roofline": {"model": "SF-IO-v1", "bytes_per_step": step_bytes, "achieved": step_gbs,
                          "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": step_gbs / HBM_PEAK_GBS},
        "kernels_ms_per_step": {k: v[0] / args.steps for k, v in per.items()},
        "wall_ms_per_step": wall_ms / args.steps,
        "energy": {"KE": ke, "PE": pe},
        "cpu_baseline": None,
    }
    if args.ensemble:
        out["ensemble"] = ensemble(sf, ic)
    if not args.no_newton:
        s.close()
        out["newton_gmres"] = newton_gmres()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(U, S, pert, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
