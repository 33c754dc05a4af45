"""params.dat and .params files (Parameters.cpp:55-111, OSUtils.cpp:84-95),
written and read by the facade (tests/cpp/params_demo.cpp), CPU only.

DumpParameters writes one "<expression> <value>" line per parameter, floats
at 30 significant digits of the FP32 value, booleans as 0/1, in the
reference's order; LoadParameters checks the grid and EvolveBackground
against the file (throwing std::string on a mismatch) and takes L3 from it;
SaveValueToFile/LoadValueFromFile round-trip a value through text."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "_build", "params_demo")


def f30(x):
    """A float written by an ostream at precision 30 (%.30g of the FP32 value)."""
    return "%.30g" % float(np.float32(x))


def test_params_files(tmp_path):
    if not os.path.exists(BIN):
        pytest.skip("params_demo not built")
    r = subprocess.run([BIN], cwd=str(tmp_path), capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    L1 = np.float32(2 * 3.14159265358979 / 0.44)
    expected = ["gridParams.N1 384", "gridParams.N2 192", "gridParams.N3 384",
                "(gridParams.dimensionality == Dimensionality::ThreeDimensional) 1",
                "(gridParams.dimensionality == Dimensionality::TwoAndAHalf) 0",
                "flowParams.L1 " + f30(L1), "flowParams.L2 " + f30(L1 / np.float32(2)), "flowParams.L3 10",
                "flowParams.Re 4000", "flowParams.Ri " + f30(0.16), "flowParams.Pr 1",
                "flowParams.EvolveBackground 0"]
    got = (tmp_path / "params.dat").read_text().splitlines()
    assert got == expected
    assert (tmp_path / "v.params").read_text() == f30(0.123456789)
    out = dict(line.split(" ", 1) for line in r.stdout.splitlines() if line.split(" ", 1)[0] in
               ("loaded", "L3", "mismatch"))
    assert np.float32(float(out["loaded"])) == np.float32(0.123456789)  # 9 digits identify an FP32 value
    assert out["L3"] == "10"
    assert out["mismatch"] == "Error: parameter gridParams.N1 not correct"
    # PrintParameters: the same lines without the TwoAndAHalf flag, at the
    # precision std::cout holds (the demo set 9 digits before)
    printed = r.stdout.splitlines()[3:]
    assert printed[:4] == expected[:4] and printed[4] == "flowParams.L1 %.9g" % float(L1)
    assert printed[-1] == "flowParams.EvolveBackground 0" and len(printed) == 11
