"""Shared set-up for parity tests: identical seeded inputs for the GPU path and the oracle."""
import numpy as np

import oracle
import stratiflow_amd as sf
from stratiflow_amd import ic

LX = 2 * np.pi / 0.44


def make_pair(N1, N2, N3, dim, Re=1000.0, Ri=0.1, Pr=1.0, dt=0.01, seed=1, gpu_prec=4, oracle_prec=8,
              amplitude=1e-3, stratification=None, project=True):
    L1, L2, L3 = LX, LX / 2, 10.0
    s = sf.IMEXRK(N1, N2, N3, dim, L1, L2, L3, Re, Ri, Pr, precision=gpu_prec)
    o = oracle.Oracle(N1, N2, N3, dim, L1, L2, L3, Re, Ri, Pr, precision=oracle_prec)
    z, zf = s.vertical_points()
    U = ic.tanh_background(zf)
    S = None if stratification is None else stratification(zf)
    s.SetBackground(U, S)
    o.set_background(U, S)
    pert = ic.perturbation(N1, N2, N3, dim, z, zf, seed, amplitude=amplitude)
    for name, arr in pert.items():
        s.set(name, arr)
        o.set_modal(name, arr)
    if project:
        s.FilterAll()
        s.PopulateNodalVariables()
        s.RemoveDivergence(0.0)
        o.filter_all()
        o.populate_nodal()
        o.remove_divergence(0.0)
    s.deltaT = dt
    o.set_dt(dt)
    return s, o


def rel_l2(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    nb = np.linalg.norm(b.ravel())
    return np.linalg.norm((a - b).ravel()) / (nb if nb > 0 else 1.0)


def fields(dim):
    return ["u1", "u2", "u3", "b", "p"] if dim == 0 else ["u1", "u3", "b", "p"]
