"""ctypes wrapper of the CPU oracle (``sf_oracle.cpp``).

TEST INFRASTRUCTURE ONLY: this is the checker for the HIP path.  Only
``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg
may import it.  The product (``stratiflow_amd``) never does.

Arrays use the reference layout (Field.h:399-410): modal fields are complex
``(N2, M1, N3)`` arrays (index ``(M1*j2 + j1)*N3 + j3``), nodal fields are real
``(N2, N1, N3)`` arrays.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libsforacle.so")

THREE_D, TWO_D, TWO_AND_A_HALF = 0, 1, 2
NEUMANN, DIRICHLET = 0, 1
MODAL_IDS = {"u1": 0, "u2": 1, "u3": 2, "b": 3, "p": 4, "r1": 5, "r2": 6, "r3": 7, "rB": 8,
             "u1t": 9, "u2t": 10, "u3t": 11, "bt": 12, "q": 13, "div": 14}
NODAL_IDS = {"U1": 0, "U2": 1, "U3": 2, "B": 3}
OPS = {"dz_ND": 0, "dz_DN": 1, "ifull": 2, "itilde": 3, "ibar": 4, "idir": 5, "dzz_N": 6, "dzz_D": 7}


def build(force=False):
    """Compile sf_oracle.cpp with g++ (no GPU toolchain needed)."""
    src = os.path.join(_HERE, "sf_oracle.cpp")
    if not force and os.path.exists(LIB_PATH) and os.path.getmtime(LIB_PATH) >= os.path.getmtime(src):
        return LIB_PATH
    subprocess.check_call(["make", "-s", "-C", _HERE])
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle library not built: run `make -C oracle` or __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        vp, i, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        dp = ctypes.POINTER(ctypes.c_double)
        L.sfo_create.restype = vp
        L.sfo_create.argtypes = [i, i, i, i, d, d, d, d, d, d, i]
        L.sfo_destroy.argtypes = [vp]
        L.sfo_set_threads.argtypes = [i]
        L.sfo_set_dt.argtypes = [vp, d]
        L.sfo_set_flow.argtypes = [vp, d, d, d]
        L.sfo_set_background.argtypes = [vp, dp, dp]
        for n in ("sfo_set_modal", "sfo_get_modal", "sfo_set_nodal", "sfo_get_nodal"):
            getattr(L, n).argtypes = [vp, i, dp]
            getattr(L, n).restype = i
        for n in ("sfo_filter_all", "sfo_populate_nodal", "sfo_set_linear_background"):
            getattr(L, n).argtypes = [vp]
        L.sfo_remove_divergence.argtypes = [vp, d]
        L.sfo_stage.argtypes = [vp, i, i]
        L.sfo_step.argtypes = [vp, i]
        L.sfo_step_linear.argtypes = [vp, i]
        L.sfo_energy.argtypes = [vp, dp, dp]
        L.sfo_inner_prod.argtypes = [vp, i, dp, dp]
        L.sfo_inner_prod.restype = d
        L.sfo_to_modal.argtypes = [vp, dp, dp]
        L.sfo_to_nodal.argtypes = [vp, dp, dp]
        L.sfo_grid.argtypes = [vp, dp, dp, dp, dp]
        L.sfo_tri_solve.argtypes = [vp, i, i, i, i, dp, dp, dp, dp, dp]
        L.sfo_apply_op.argtypes = [vp, i, dp, dp]
        L.sfo_apply_op.restype = i
        L.sfo_prepare_adjoint.argtypes = [vp]
        L.sfo_jover_k.argtypes = [vp]
        L.sfo_jover_k.restype = d
        L.sfo_cfl.argtypes = [vp, i]
        L.sfo_cfl.restype = d
        L.sfo_set_evolve_background.argtypes = [vp, i]
        L.sfo_get_background.argtypes = [vp, dp]
        L.sfo_adjoint.argtypes = [vp, i, ctypes.POINTER(dp), ctypes.POINTER(dp), d]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def set_threads(n):
    lib().sfo_set_threads(int(n))


class Oracle:
    """One restated IMEXRK solver instance (float32 or float64 arithmetic)."""

    def __init__(self, N1, N2, N3, dim, L1, L2, L3, Re, Ri, Pr, precision=8):
        self.N1, self.N2, self.N3 = N1, N2, N3
        self.M1 = N1 // 2 + 1
        self.dim = dim
        self.precision = precision
        self.h = lib().sfo_create(N1, N2, N3, dim, L1, L2, L3, Re, Ri, Pr, precision)
        if not self.h:
            raise RuntimeError("sfo_create failed")
        self.dt = 0.01  # the solver's initial deltaT

    def __del__(self):
        if getattr(self, "h", None):
            lib().sfo_destroy(self.h)
            self.h = None

    @property
    def modal_shape(self):
        return (self.N2, self.M1, self.N3)

    @property
    def nodal_shape(self):
        return (self.N2, self.N1, self.N3)

    def set_dt(self, dt):
        self.dt = float(dt)
        lib().sfo_set_dt(self.h, dt)

    def set_flow(self, Re, Ri, Pr):
        lib().sfo_set_flow(self.h, Re, Ri, Pr)

    def set_background(self, U, S=None):
        U = np.ascontiguousarray(U, dtype=np.float64)
        Sp = None if S is None else _p(np.ascontiguousarray(S, dtype=np.float64))
        if S is not None:
            S = np.ascontiguousarray(S, dtype=np.float64)
            Sp = _p(S)
        lib().sfo_set_background(self.h, _p(U), Sp)

    def set_modal(self, name, arr):
        a = np.ascontiguousarray(np.asarray(arr, dtype=np.complex128).reshape(self.modal_shape))
        if lib().sfo_set_modal(self.h, MODAL_IDS[name], _p(a.view(np.float64))) != 0:
            raise KeyError(name)

    def get_modal(self, name):
        a = np.zeros(self.modal_shape, dtype=np.complex128)
        if lib().sfo_get_modal(self.h, MODAL_IDS[name], _p(a.view(np.float64))) != 0:
            raise KeyError(name)
        return a

    def set_nodal(self, name, arr):
        a = np.ascontiguousarray(np.asarray(arr, dtype=np.float64).reshape(self.nodal_shape))
        lib().sfo_set_nodal(self.h, NODAL_IDS[name], _p(a))

    def get_nodal(self, name):
        a = np.zeros(self.nodal_shape, dtype=np.float64)
        lib().sfo_get_nodal(self.h, NODAL_IDS[name], _p(a))
        return a

    def filter_all(self):
        lib().sfo_filter_all(self.h)

    def populate_nodal(self):
        lib().sfo_populate_nodal(self.h)

    def remove_divergence(self, mult):
        lib().sfo_remove_divergence(self.h, mult)

    def set_linear_background(self):
        lib().sfo_set_linear_background(self.h)

    def stage(self, k, linear=False):
        lib().sfo_stage(self.h, k, int(linear))

    def step(self, n=1):
        lib().sfo_step(self.h, n)

    def step_linear(self, n=1):
        lib().sfo_step_linear(self.h, n)

    def jover_k(self):
        """IMEXRK::JoverK (IMEXRK.h:445-459) from the nodal B and the modal u3."""
        return lib().sfo_jover_k(self.h)

    def set_evolve_background(self, on=True):
        """FlowParams::EvolveBackground: U_ takes a Crank-Nicolson diffusion solve per
        stage (IMEXRK.cpp:100-104)."""
        lib().sfo_set_evolve_background(self.h, int(bool(on)))

    def background(self):
        """The current background velocity U_ (N3 values)."""
        U = np.zeros(self.N3)
        lib().sfo_get_background(self.h, _p(U))
        return U

    def cfl(self, linear=False):
        """CFL (IMEXRK.h:355-374) or CFLlinear (IMEXRK.h:376-393) for the current dt,
        without their dt update; CFL reads the nodal U1..U3 as they are, CFLlinear
        the nodal linearisation state of set_linear_background."""
        return lib().sfo_cfl(self.h, int(bool(linear)))

    def prepare_adjoint(self):
        """PrepareRunAdjoint (IMEXRK.h:183-199): zero pressure, populate nodal fields."""
        lib().sfo_prepare_adjoint(self.h)

    def step_adjoint(self, below, above):
        """One TimeStepAdjoint (IMEXRK.h:86-126).  below/above: dicts u1, u2, u3, b of
        modal arrays (reference layout); a missing field is zero."""
        self._adjoint(0, below, above, 0.0)

    def rhs_adjoint(self, below, above, frac):
        """Explicit terms of BuildRHSAdjoint (IMEXRK.cpp:260-379) into r1..rB, with the
        direct state interpolated at frac and the nodal fields as they are."""
        self._adjoint(1, below, above, frac)

    def rhs_linear(self):
        """Explicit terms of BuildRHSLinear (IMEXRK.cpp:195-258) into r1..rB."""
        self._adjoint(2, {}, {}, 0.0)

    def _adjoint(self, mode, below, above, frac):
        keep = []

        def ptrs(st):
            arr = (ctypes.POINTER(ctypes.c_double) * 4)()
            for i, n in enumerate(("u1", "u2", "u3", "b")):
                if n in st and st[n] is not None:
                    a = np.ascontiguousarray(np.asarray(st[n], dtype=np.complex128).reshape(self.modal_shape))
                    keep.append(a)
                    arr[i] = _p(a.view(np.float64))
                else:
                    arr[i] = ctypes.POINTER(ctypes.c_double)()
            return arr
        lib().sfo_adjoint(self.h, mode, ptrs(below), ptrs(above), frac)

    def energy(self):
        ke, pe = ctypes.c_double(), ctypes.c_double()
        lib().sfo_energy(self.h, ctypes.byref(ke), ctypes.byref(pe))
        return ke.value, pe.value

    def inner_prod(self, a, b, bc):
        a = np.ascontiguousarray(np.asarray(a, dtype=np.complex128).reshape(self.modal_shape))
        b = np.ascontiguousarray(np.asarray(b, dtype=np.complex128).reshape(self.modal_shape))
        return lib().sfo_inner_prod(self.h, bc, _p(a.view(np.float64)), _p(b.view(np.float64)))

    def to_modal(self, nodal):
        x = np.ascontiguousarray(np.asarray(nodal, dtype=np.float64).reshape(self.nodal_shape))
        out = np.zeros(self.modal_shape, dtype=np.complex128)
        lib().sfo_to_modal(self.h, _p(x), _p(out.view(np.float64)))
        return out

    def to_nodal(self, modal):
        x = np.ascontiguousarray(np.asarray(modal, dtype=np.complex128).reshape(self.modal_shape))
        out = np.zeros(self.nodal_shape, dtype=np.float64)
        lib().sfo_to_nodal(self.h, _p(x.view(np.float64)), _p(out))
        return out

    def grid(self):
        z, zf, DY, DYF = (np.zeros(self.N3) for _ in range(4))
        lib().sfo_grid(self.h, _p(z), _p(zf), _p(DY), _p(DYF))
        return z, zf, DY, DYF

    def apply_op(self, name, col):
        col = np.ascontiguousarray(col, dtype=np.float64)
        out = np.zeros(self.N3)
        if lib().sfo_apply_op(self.h, OPS[name], _p(col), _p(out)) != 0:
            raise KeyError(name)
        return out

    def tri_solve(self, kind, stage, j1, j2, d):
        d = np.ascontiguousarray(np.asarray(d, dtype=np.complex128))
        x = np.zeros(self.N3, dtype=np.complex128)
        a, b, c = np.zeros(self.N3), np.zeros(self.N3), np.zeros(self.N3)
        lib().sfo_tri_solve(self.h, kind, stage, j1, j2, _p(d.view(np.float64)), _p(x.view(np.float64)),
                            _p(a), _p(b), _p(c))
        return x, (a, b, c)
