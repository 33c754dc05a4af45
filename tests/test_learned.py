"""Learned-preconditioner Newton-GMRES against plain GMRES (tests/cpp/learned_demo.cpp).

The network is trained on the operator applications of plain solves and then
used as GMRES's flexible right preconditioner on held-out problems; plain GMRES
judges it: a preconditioned solve that misses the tolerance is repeated without
the network.  The test checks that training lowers the loss and that every
Newton update equals the plain solver's, and prints the cost of both methods
(operator applications, wall time) -- what the trained network gains or loses.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "_build", "learned_demo")


@pytest.mark.gpu
def test_learned_preconditioner_judged_by_plain_gmres():
    r = subprocess.run([BIN, "3", "2", "41", "1e-2", "40", "100"], capture_output=True, text=True, timeout=280)
    assert r.returncode == 0, r.stdout + r.stderr
    loss = None
    tests = []
    for line in r.stdout.splitlines():
        t = line.split()
        if t[0] == "loss":
            loss = (float(t[1]), float(t[2]), int(t[3]))
        elif t[0] == "test":
            tests.append(dict(plain_mv=int(t[2]), plain_s=float(t[3]), learned_mv=int(t[4]), learned_s=float(t[5]),
                              plain_res=float(t[6]), learned_res=float(t[7]), diff=float(t[8])))
    assert loss is not None and loss[1] < loss[0], loss
    assert len(tests) == 2
    for k, t in enumerate(tests):
        print("problem %d: plain %d matvecs %.3f s (res %.2e); learned %d matvecs %.3f s (res %.2e); "
              "time ratio plain/learned %.2f" % (k, t["plain_mv"], t["plain_s"], t["plain_res"], t["learned_mv"],
                                                t["learned_s"], t["learned_res"], t["plain_s"] / t["learned_s"]))
        # the judged update equals the plain one: exactly after a fallback,
        # within the GMRES tolerance when the preconditioned solve converged
        assert t["diff"] <= (0.0 if t["learned_res"] >= 1e-2 else 0.05), t


@pytest.mark.gpu
def test_network_correcting_the_analytic_base():
    """With the advected diffusion preconditioner as its base (learned_demo's
    ninth argument) the trained network's preconditioned solves converge on
    held-out problems without falling back and with fewer operator
    applications than plain GMRES."""
    r = subprocess.run([BIN, "3", "2", "61", "2e-2", "40", "100", "1", "0", "0.3"], capture_output=True, text=True,
                       timeout=280)
    assert r.returncode == 0, r.stdout + r.stderr
    tests = [line.split() for line in r.stdout.splitlines() if line.startswith("test")]
    assert len(tests) == 2
    for t in tests:
        plain_mv, learned_mv, learned_res, fallbacks = int(t[2]), int(t[4]), float(t[7]), int(t[9])
        print("plain %d, learned with base %d (res %.2e, fallbacks %d)" % (plain_mv, learned_mv, learned_res,
                                                                         fallbacks), flush=True)
        assert fallbacks == 0 and learned_res < 2e-2
        assert learned_mv < plain_mv
