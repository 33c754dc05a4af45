"""Write the reference energy series of the 3D throughput workload.

384x192x384, Re=4000, Ri=0.16, Pr=1, L1 = 2*pi/0.44, L2 = L1/2, L3 = 10,
U = tanh z and background buoyancy tanh z (dB/dz = sech^2 z at the Neumann
points), the seeded perturbation of amplitude 1e-3 (seed 20261202), dt = 0.005:
the CPU oracle in FP32 and in FP64 runs FullEvolve's start-up (FilterAll,
PopulateNodalVariables, RemoveDivergence(0)) and 200 steps, recording KE and
PE (IMEXRK.h:395-411) before the first step and after every step.

Output: tests/golden/c2_series_384x192x384.npz with ke32, pe32, ke64, pe64
(201 values each).  The oracle is the checker; this script runs in the build
container (hours of CPU time: run FP32 and FP64 in two processes with
--precision) and its output is committed as a fixture.

usage: python tests/golden/make_c2_series.py --precision {4,8} --threads N
       python tests/golden/make_c2_series.py --merge
"""
import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

N1, N2, N3 = 384, 192, 384
RE, RI, PR, L3, DT, STEPS, SEED = 4000.0, 0.16, 1.0, 10.0, 0.005, 200, 20261202
OUT = os.path.join(HERE, "c2_series_384x192x384.npz")


def run(precision, threads):
    import oracle
    from stratiflow_amd import ic
    oracle.set_threads(threads)
    L1 = 2 * np.pi / 0.44
    o = oracle.Oracle(N1, N2, N3, oracle.THREE_D, L1, L1 / 2, L3, RE, RI, PR, precision=precision)
    z, zf, _, _ = o.grid()
    o.set_background(ic.tanh_background(zf), ic.tanh_stratification(zf))
    for n, a in ic.perturbation(N1, N2, N3, oracle.THREE_D, z, zf, SEED, amplitude=1e-3).items():
        o.set_modal(n, a)
    o.filter_all()
    o.populate_nodal()
    o.remove_divergence(0.0)
    o.set_dt(DT)
    ke, pe = [], []
    k, p = o.energy()
    ke.append(k)
    pe.append(p)
    t0 = time.time()
    for step in range(STEPS):
        o.step(1)
        k, p = o.energy()
        ke.append(k)
        pe.append(p)
        print("fp%d step %d ke %.12e pe %.12e (%.0f s)" % (8 * precision, step + 1, k, p, time.time() - t0),
              flush=True)
    np.savez(OUT.replace(".npz", "_fp%d.npz" % (8 * precision)), ke=np.array(ke), pe=np.array(pe))


def merge():
    a = np.load(OUT.replace(".npz", "_fp32.npz"))
    b = np.load(OUT.replace(".npz", "_fp64.npz"))
    np.savez(OUT, ke32=a["ke"], pe32=a["pe"], ke64=b["ke"], pe64=b["pe"],
             meta=np.array("384x192x384 Re=4000 Ri=0.16 Pr=1 tanh U and B dt=0.005 seed=20261202 amp=1e-3"))
    for p in ("_fp32.npz", "_fp64.npz"):
        os.remove(OUT.replace(".npz", p))


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--precision", type=int, choices=(4, 8))
    ap.add_argument("--threads", type=int, default=4)
    ap.add_argument("--merge", action="store_true")
    a = ap.parse_args()
    if a.merge:
        merge()
    else:
        run(a.precision, a.threads)
