"""Regenerate the golden fixtures from the FP64 CPU oracle.

The reference has no fixtures of its own (no tests directory), so these are
the oracle's FP64 outputs on small seeded cases.  They pin the GPU path
independently of a live oracle build and guard the oracle against regressions.
Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from stratiflow_amd import ic  # noqa: E402

LX = 2 * np.pi / 0.44
CASES = {
    "c2d_12x1x16": dict(N1=12, N2=1, N3=16, dim=oracle.TWO_D),
    "c3d_12x6x16": dict(N1=12, N2=6, N3=16, dim=oracle.THREE_D),
    "c2d_48x1x64": dict(N1=48, N2=1, N3=64, dim=oracle.TWO_D),
}
FLOW = dict(Re=800.0, Ri=0.12, Pr=0.9)
DT = 0.02
SEED = 7
AMP = 0.1


def vertical_points(N3):
    o = oracle.Oracle(12, 1, N3, oracle.TWO_D, LX, LX / 2, 10.0, 1, 1, 1)
    z, zf, _, _ = o.grid()
    return z, zf


def run_case(N1, N2, N3, dim):
    z, zf = vertical_points(N3)
    U = ic.tanh_background(zf)
    S = ic.tanh_stratification(zf)
    pert = ic.perturbation(N1, N2, N3, dim, z, zf, SEED, amplitude=AMP)
    out = {"U": U, "S": S, "z": z, "zf": zf}
    for n, a in pert.items():
        out["in_" + n] = a
    names = ["u1", "u2", "u3", "b", "p"] if dim == oracle.THREE_D else ["u1", "u3", "b", "p"]

    def fresh():
        o = oracle.Oracle(N1, N2, N3, dim, LX, LX / 2, 10.0, FLOW["Re"], FLOW["Ri"], FLOW["Pr"], precision=8)
        o.set_background(U, S)
        for n, a in pert.items():
            o.set_modal(n, a)
        o.filter_all()
        o.populate_nodal()
        o.remove_divergence(0.0)
        o.set_dt(DT)
        return o

    o = fresh()
    for n in names:
        out["proj_" + n] = o.get_modal(n)
    o.stage(0)
    for n in names + ["r1", "r3", "rB"] + (["r2"] if dim == oracle.THREE_D else []):
        out["stage0_" + n] = o.get_modal(n)
    o = fresh()
    o.step(1)
    for n in names:
        out["step1_" + n] = o.get_modal(n)
    out["step1_energy"] = np.array(o.energy())
    o.step(9)
    for n in names:
        out["step10_" + n] = o.get_modal(n)
    out["step10_energy"] = np.array(o.energy())
    out["nodal_U1_step10"] = o.get_nodal("U1")
    return out


def main():
    for name, c in CASES.items():
        data = run_case(**c)
        meta = dict(c, **FLOW, dt=DT, seed=SEED, amplitude=AMP, L1=LX, L2=LX / 2, L3=10.0)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), meta=np.array(repr(meta)), **data)
        print("wrote", name)


if __name__ == "__main__":
    main()
