"""The reference's own drivers compile and link unchanged against the facade.

The reference's callers of the hot path (ExtendedStateVector.h,
HopfBifurcation.h, PseudoArclength.h and the mains Adjoint.cpp, Direct.cpp,
FindPeriodic.cpp, FieldNorm.cpp, Stability3D.cpp, TrackSolution.cpp,
FindHopf.cpp, FindCriticalPoint.cpp, NewtonKrylov.cpp, PseudoArclength.cpp)
include "StateVector.h", "NewtonKrylov.h", "BasicNewtonKrylov.h" and
"Arnoldi.h".  A scratch directory holds one-line headers of those names that
include the facade (include/stratiflow/), next to symbolic links to the
callers where they lie in the reference tree: a quoted include resolves in the
directory of the including file's path, so the callers pick up the facade
instead of the reference's own headers.  Nothing of the reference is copied.

InverseIteration.cpp is left out: it uses Eigen's MatrixX/JacobiSVD directly
(InverseIteration.cpp:42-101), and Eigen is not part of the boundary.

Skipped where the reference tree is absent.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
INCLUDE = os.path.join(ROOT, "include")
LIBDIR = os.path.join(ROOT, "stratiflow_amd")

SHIMS = {
    "StateVector.h": "stratiflow/stratiflow.h",
    "NewtonKrylov.h": "stratiflow/NewtonKrylov.h",
    "BasicNewtonKrylov.h": "stratiflow/BasicNewtonKrylov.h",
    "Arnoldi.h": "stratiflow/Arnoldi.h",
}
HEADERS = ["ExtendedStateVector.h", "HopfBifurcation.h", "PseudoArclength.h"]
MAINS = ["Adjoint.cpp", "Direct.cpp", "FindPeriodic.cpp", "FieldNorm.cpp", "Stability3D.cpp", "TrackSolution.cpp",
         "FindHopf.cpp", "FindCriticalPoint.cpp", "NewtonKrylov.cpp", "PseudoArclength.cpp"]

pytestmark = pytest.mark.skipif(not os.path.isdir(REF) or shutil.which("g++") is None,
                                reason="needs the reference tree and g++")


@pytest.fixture(scope="module")
def shimdir(tmp_path_factory):
    d = tmp_path_factory.mktemp("dropin")
    for name, target in SHIMS.items():
        (d / name).write_text("#pragma once\n#include <%s>\n" % target)
    for name in HEADERS + MAINS:
        os.symlink(os.path.join(REF, name), str(d / name))
    return d


def _gxx(args, cwd):
    r = subprocess.run(["g++", "-std=c++14", "-I" + INCLUDE] + args, cwd=cwd, stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=300)
    return r.returncode, r.stdout


@pytest.mark.parametrize("name", HEADERS)
def test_caller_header_compiles(shimdir, name):
    rc, out = _gxx(["-fsyntax-only", "-x", "c++", str(shimdir / name)], str(shimdir))
    assert rc == 0, out[-3000:]


@pytest.mark.parametrize("name", MAINS)
def test_driver_links(shimdir, name):
    lib = os.path.join(LIBDIR, "libsfhip.so")
    if not os.path.exists(lib):
        pytest.skip("libsfhip.so not built")
    exe = str(shimdir / (name[:-4] + ".bin"))
    rc, out = _gxx(["-O0", "-o", exe, str(shimdir / name), "-L" + LIBDIR, "-lsfhip"], str(shimdir))
    assert rc == 0, out[-3000:]
    assert os.path.exists(exe)
