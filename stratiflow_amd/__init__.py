"""Host-side mirror of Stratiflow's IMEXRK / Field interface over ``libsfhip.so``.

The product is the gfx950 library ``libsfhip.so`` with the C ABI declared in
``include/sfhip.h``.  This module binds that ABI with ctypes so that Python
callers (the test-suite, ``bench.py``) drive the same entry points a C++
caller uses through ``include/stratiflow/``.  There is no fallback: if the
library is missing or no GPU is present, every call raises.

Names follow the reference: ``IMEXRK`` (IMEXRK.h:31-670) with ``TimeStep``,
``RemoveDivergence``, ``FilterAll``, ``PopulateNodalVariables``, ``KE``,
``PE``, ``SetBackground``, ``deltaT``; fields follow ``NeumannModal`` /
``DirichletModal`` / ``NeumannNodal`` / ``DirichletNodal`` (Stratiflow.h:8-49).
Host arrays use the reference layouts (Field.h:399-410): modal fields are
complex ``(N2, N1//2+1, N3)``, nodal fields real ``(N2, N1, N3)``.
"""
import ctypes
import os
import weakref

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsfhip.so")

THREE_D, TWO_D, TWO_AND_A_HALF = 0, 1, 2
BATCH_GROUPED, BATCH_MIXED, BATCH_MIXED_R8, BATCH_MIXED8 = 0, 1, 2, 2
XPASS_PLANS = {0: "grouped", 1: "mixed", 2: "mixed_r8"}
MODAL_PLANS = {0: "grouped", 1: "mixed4", 2: "mixed8"}
NEUMANN, DIRICHLET = 0, 1
NODAL, MODAL, ONE_D = 0, 1, 2
STATE_IDS = {"u1": 0, "u2": 1, "u3": 2, "b": 3, "p": 4, "U1": 5, "U2": 6, "U3": 7, "B": 8,
             "r1": 9, "r2": 10, "r3": 11, "rB": 12, "u1_tot": 13, "u2_tot": 14, "u3_tot": 15, "b_tot": 16}


class SFError(RuntimeError):
    """Non-zero status from the C ABI (message from sf_last_error)."""

    def __init__(self, code, msg):
        super().__init__("sfhip error %d: %s" % (code, msg))
        self.code = code


class _Grid(ctypes.Structure):
    _fields_ = [("N1", ctypes.c_int), ("N2", ctypes.c_int), ("N3", ctypes.c_int), ("dimensionality", ctypes.c_int)]


class _Flow(ctypes.Structure):
    _fields_ = [("L1", ctypes.c_double), ("L2", ctypes.c_double), ("L3", ctypes.c_double),
                ("Re", ctypes.c_double), ("Ri", ctypes.c_double), ("Pr", ctypes.c_double),
                ("evolve_background", ctypes.c_int)]


_lib = None

# every extern "C" symbol of include/sfhip.h
EXPORTED = [
    "sf_ctx_create", "sf_ctx_destroy", "sf_last_error", "sf_set_dt", "sf_get_dt", "sf_set_flow",
    "sf_set_background", "sf_vertical_points", "sf_field_alloc", "sf_field_free", "sf_field_kind", "sf_field_upload",
    "sf_field_download", "sf_field_copy", "sf_field_zero", "sf_field_scale", "sf_field_axpby",
    "sf_field_zero_ends", "sf_field_neumann_ends", "sf_field_filter", "sf_to_modal", "sf_to_nodal",
    "sf_ddx", "sf_ddy", "sf_ddz", "sf_field_diffusion_precondition", "sf_inner_prod", "sf_state_field", "sf_step", "sf_step_linear",
    "sf_set_linear_background", "sf_step_adjoint", "sf_jover_k", "sf_field_add_profile", "sf_field_mean_profile",
    "sf_dot_f64", "sf_field_enforce_bcs", "sf_phase_shift", "sf_remove_average", "sf_state_bind", "sf_state_fetch",
    "sf_energy_f64", "sf_stage", "sf_remove_divergence", "sf_filter_all", "sf_populate_nodal",
    "sf_energy", "sf_cfl", "sf_cfl_linear", "sf_cfl_linear_noalloc", "sf_refresh_nodal", "sf_get_background", "sf_synchronize", "sf_time_steps", "sf_time_steps_detailed",
    "sf_batch_create", "sf_batch_destroy", "sf_batch_last_error", "sf_batch_step", "sf_batch_time_steps",
    "sf_batch_info", "sf_batch_plan", "sf_batch_select",
    "sf_net_create", "sf_net_destroy", "sf_net_last_error", "sf_net_apply", "sf_net_add_pair", "sf_net_clear_pairs",
    "sf_net_samples", "sf_net_train", "sf_net_param_count", "sf_net_get_params", "sf_net_set_params",
    "sf_net_geometry", "sf_net_forward_raw", "sf_net_grad_raw", "sf_net_set_dilations",
]
KERNEL_CLASSES = ["yinv", "xpass", "yfwd", "modal"]


def lib():
    """Load libsfhip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libsfhip.so not found at %s: run `python -m stratiflow_amd.build`" % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    vp, i, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    dp = ctypes.POINTER(ctypes.c_double)
    vpp = ctypes.POINTER(ctypes.c_void_p)
    sig = {
        "sf_ctx_create": [ctypes.POINTER(_Grid), ctypes.POINTER(_Flow), i, vpp],
        "sf_ctx_destroy": [vp],
        "sf_set_dt": [vp, d],
        "sf_get_dt": [vp, dp],
        "sf_set_flow": [vp, ctypes.POINTER(_Flow)],
        "sf_set_background": [vp, dp, dp],
        "sf_vertical_points": [vp, dp, dp],
        "sf_field_alloc": [vp, i, i, vpp],
        "sf_field_free": [vp],
        "sf_field_kind": [vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)],
        "sf_field_upload": [vp, dp],
        "sf_field_download": [vp, dp],
        "sf_field_copy": [vp, vp],
        "sf_field_zero": [vp],
        "sf_field_scale": [vp, d],
        "sf_field_axpby": [vp, d, vp, d],
        "sf_field_zero_ends": [vp],
        "sf_field_neumann_ends": [vp],
        "sf_field_filter": [vp],
        "sf_to_modal": [vp, vp],
        "sf_to_nodal": [vp, vp],
        "sf_ddx": [vp, vp],
        "sf_ddy": [vp, vp],
        "sf_ddz": [vp, vp],
        "sf_field_diffusion_precondition": [vp, vp, d, d, d, d],
        "sf_inner_prod": [vp, vp, dp],
        "sf_state_field": [vp, i, vpp],
        "sf_step": [vp, i],
        "sf_step_linear": [vp, i],
        "sf_set_linear_background": [vp, vp, vp, vp, vp],
        "sf_step_adjoint": [vp, ctypes.POINTER(vp), ctypes.POINTER(vp)],
        "sf_jover_k": [vp, dp],
        "sf_field_add_profile": [vp, dp],
        "sf_field_mean_profile": [vp, dp],
        "sf_dot_f64": [vp, vp, dp],
        "sf_field_enforce_bcs": [vp],
        "sf_phase_shift": [vp, d],
        "sf_remove_average": [vp],
        "sf_state_bind": [vp, vp, vp, vp, vp, vp],
        "sf_state_fetch": [vp, vp, vp, vp, vp, vp],
        "sf_energy_f64": [vp, dp, dp],
        "sf_stage": [vp, i, i],
        "sf_remove_divergence": [vp, d],
        "sf_filter_all": [vp],
        "sf_populate_nodal": [vp],
        "sf_energy": [vp, dp, dp],
        "sf_cfl": [vp, dp],
        "sf_cfl_linear": [vp, dp],
        "sf_cfl_linear_noalloc": [vp, dp],
        "sf_refresh_nodal": [vp],
        "sf_get_background": [vp, dp],
        "sf_synchronize": [vp],
        "sf_time_steps": [vp, i, dp],
        "sf_time_steps_detailed": [vp, i, dp, dp, ctypes.POINTER(ctypes.c_int)],
        "sf_batch_create": [vpp, i, vpp],
        "sf_batch_destroy": [vp],
        "sf_batch_step": [vp, i],
        "sf_batch_time_steps": [vp, i, dp],
        "sf_batch_info": [vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                          ctypes.POINTER(ctypes.c_int)],
        "sf_batch_plan": [vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int), dp],
        "sf_batch_select": [vp, i, i],
        "sf_net_create": [vp, i, i, ctypes.c_ulonglong, vpp],
        "sf_net_destroy": [vp],
        "sf_net_apply": [vp, vpp, vpp],
        "sf_net_add_pair": [vp, vpp, vpp],
        "sf_net_clear_pairs": [vp],
        "sf_net_samples": [vp, ctypes.POINTER(ctypes.c_int)],
        "sf_net_train": [vp, i, i, d, dp, dp],
        "sf_net_param_count": [vp, ctypes.POINTER(ctypes.c_size_t)],
        "sf_net_get_params": [vp, ctypes.POINTER(ctypes.c_float)],
        "sf_net_set_params": [vp, ctypes.POINTER(ctypes.c_float)],
        "sf_net_set_dilations": [vp, ctypes.POINTER(ctypes.c_int)],
        "sf_net_geometry": [vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                            ctypes.POINTER(ctypes.c_int)],
        "sf_net_forward_raw": [vp, ctypes.POINTER(ctypes.c_float), i, ctypes.POINTER(ctypes.c_float)],
        "sf_net_grad_raw": [vp, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float), i,
                            ctypes.POINTER(ctypes.c_float), dp],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = ctypes.c_int
    L.sf_last_error.argtypes = [vp]
    L.sf_last_error.restype = ctypes.c_char_p
    L.sf_batch_last_error.argtypes = [vp]
    L.sf_batch_last_error.restype = ctypes.c_char_p
    L.sf_net_last_error.argtypes = [vp]
    L.sf_net_last_error.restype = ctypes.c_char_p
    _lib = L
    return L


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Field:
    """A device field (NeumannModal, DirichletNodal, ...), owned or borrowed."""

    def __init__(self, solver, handle, kind, bc, owned):
        self.solver = solver
        self.h = handle
        self.kind = kind
        self.bc = bc
        self.owned = owned

    def close(self):
        """Free an owned field (sf_field_free); the solver does this for its
        fields before it is closed."""
        if self.owned and getattr(self, "h", None):
            lib().sf_field_free(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def shape(self):
        s = self.solver
        if self.kind == MODAL:
            return (s.N2, s.M1, s.N3)
        if self.kind == NODAL:
            return (s.N2, s.N1, s.N3)
        return (s.N3,)

    def upload(self, arr):
        if self.kind == MODAL:
            a = np.ascontiguousarray(np.asarray(arr, dtype=np.complex128).reshape(self.shape))
            self.solver._chk(lib().sf_field_upload(self.h, _p(a.view(np.float64))))
        else:
            a = np.ascontiguousarray(np.asarray(arr, dtype=np.float64).reshape(self.shape))
            self.solver._chk(lib().sf_field_upload(self.h, _p(a)))

    def download(self):
        if self.kind == MODAL:
            a = np.zeros(self.shape, dtype=np.complex128)
            self.solver._chk(lib().sf_field_download(self.h, _p(a.view(np.float64))))
        else:
            a = np.zeros(self.shape, dtype=np.float64)
            self.solver._chk(lib().sf_field_download(self.h, _p(a)))
        return a

    # Field / ModalField / NodalField operations
    def Zero(self):
        self.solver._chk(lib().sf_field_zero(self.h))

    def PhaseShift(self, shift):
        """ModalField::PhaseShift (Field.h:1168-1174)."""
        self.solver._chk(lib().sf_phase_shift(self.h, float(shift)))

    def RemoveAverage(self):
        """RemoveAverage (Integration.h:72-75)."""
        self.solver._chk(lib().sf_remove_average(self.h))

    def EnforceBCs(self):
        """ZeroEnds (Dirichlet) or NeumannEnds (Neumann), StateVector.h:415-421."""
        self.solver._chk(lib().sf_field_enforce_bcs(self.h))

    def add_profile(self, prof):
        """NodalField += Nodal1D (N3 values broadcast over every stack)."""
        a = np.ascontiguousarray(prof, dtype=np.float64)
        self.solver._chk(lib().sf_field_add_profile(self.h, _p(a)))

    def mean_profile(self):
        """HorizontalAverage (Integration.h:37-42): Re of the (0,0) stack of a modal field."""
        out = np.zeros(self.solver.N3)
        self.solver._chk(lib().sf_field_mean_profile(self.h, _p(out)))
        return out

    def assign(self, other):
        self.solver._chk(lib().sf_field_copy(self.h, other.h))

    def __imul__(self, a):
        self.solver._chk(lib().sf_field_scale(self.h, float(a)))
        return self

    def axpby(self, a, x, b):
        self.solver._chk(lib().sf_field_axpby(self.h, float(a), x.h if x is not None else None, float(b)))

    def MulAdd(self, a, x):
        self.axpby(a, x, 1.0)

    def ZeroEnds(self):
        self.solver._chk(lib().sf_field_zero_ends(self.h))

    def NeumannEnds(self):
        self.solver._chk(lib().sf_field_neumann_ends(self.h))

    def Filter(self):
        self.solver._chk(lib().sf_field_filter(self.h))

    def diffusion_precondition(self, out, tnu, delta, kmax=0.0, tadv=0.0):
        """out = (tnu*(Dzz - k^2) - i*tadv*kx*U(z) - delta)^-1 ZeroEnds(self) - self/2
        on modes with |k| <= kmax, -self elsewhere (sf_field_diffusion_precondition)."""
        self.solver._chk(lib().sf_field_diffusion_precondition(self.h, out.h, float(tnu), float(delta), float(kmax),
                                                               float(tadv)))

    def ToModal(self, other):
        self.solver._chk(lib().sf_to_modal(self.h, other.h))

    def ToNodal(self, other):
        self.solver._chk(lib().sf_to_nodal(self.h, other.h))


class IMEXRK:
    """The IMEX-RK3/CN solver of IMEXRK.h:31-670 on one GPU.

    ``precision`` 4 is the reference's default ``stratifloat`` (float), 8 the
    ``-DDOUBLE=On`` build.
    """

    def __init__(self, N1, N2, N3, dimensionality, L1, L2, L3, Re, Ri, Pr, precision=4, evolve_background=False):
        L = lib()
        self.N1, self.N2, self.N3 = N1, N2, N3
        self.M1 = N1 // 2 + 1
        self.dimensionality = dimensionality
        self.precision = precision
        self.flow = dict(L1=L1, L2=L2, L3=L3, Re=Re, Ri=Ri, Pr=Pr, EvolveBackground=bool(evolve_background))
        g = _Grid(N1, N2, N3, dimensionality)
        f = _Flow(L1, L2, L3, Re, Ri, Pr, int(bool(evolve_background)))
        h = ctypes.c_void_p()
        rc = L.sf_ctx_create(ctypes.byref(g), ctypes.byref(f), precision, ctypes.byref(h))
        if rc != 0:
            raise SFError(rc, L.sf_last_error(None).decode())
        self.h = h
        self._state = {}
        # owned fields, networks and batches made from this context: the C ABI
        # requires them to be freed before the context is destroyed
        self._deps = weakref.WeakSet()

    def close(self):
        if getattr(self, "h", None):
            for d in list(self._deps):
                d.close()
            self._state = {}
            lib().sf_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        if rc != 0:
            raise SFError(rc, lib().sf_last_error(self.h).decode())

    # ---- parameters ----
    @property
    def deltaT(self):
        d = ctypes.c_double()
        self._chk(lib().sf_get_dt(self.h, ctypes.byref(d)))
        return d.value

    @deltaT.setter
    def deltaT(self, dt):
        """IMEXRK::deltaT followed by UpdateForTimestep (IMEXRK.h:502-545)."""
        self._chk(lib().sf_set_dt(self.h, float(dt)))

    def UpdateForTimestep(self):
        self.deltaT = self.deltaT

    def ResetForParams(self, **flow):
        self.flow.update(flow)
        f = _Flow(self.flow["L1"], self.flow["L2"], self.flow["L3"], self.flow["Re"], self.flow["Ri"],
                  self.flow["Pr"], int(bool(self.flow["EvolveBackground"])))
        self._chk(lib().sf_set_flow(self.h, ctypes.byref(f)))

    def Background(self):
        """The background velocity U_ (N3 values) as the solver holds it now: it
        changes every stage when EvolveBackground is set (IMEXRK.cpp:100-104)."""
        U = np.zeros(self.N3)
        self._chk(lib().sf_get_background(self.h, _p(U)))
        return U

    def vertical_points(self):
        """(z, zf): VerticalPoints / VerticalPointsFractional (Field.cpp:3-31)."""
        z, zf = np.zeros(self.N3), np.zeros(self.N3)
        self._chk(lib().sf_vertical_points(self.h, _p(z), _p(zf)))
        return z, zf

    def SetBackground(self, U, dBdz=None):
        U = np.ascontiguousarray(U, dtype=np.float64)
        S = None if dBdz is None else np.ascontiguousarray(dBdz, dtype=np.float64)
        self._chk(lib().sf_set_background(self.h, _p(U), _p(S) if S is not None else None))

    # ---- fields ----
    def field(self, kind, bc):
        h = ctypes.c_void_p()
        self._chk(lib().sf_field_alloc(self.h, kind, bc, ctypes.byref(h)))
        f = Field(self, h, kind, bc, True)
        self._deps.add(f)
        return f

    def state(self, name):
        if name not in self._state:
            h = ctypes.c_void_p()
            self._chk(lib().sf_state_field(self.h, STATE_IDS[name], ctypes.byref(h)))
            k, b = ctypes.c_int(), ctypes.c_int()
            self._chk(lib().sf_field_kind(h, ctypes.byref(k), ctypes.byref(b)))
            self._state[name] = Field(self, h, k.value, b.value, False)
        return self._state[name]

    def get(self, name):
        return self.state(name).download()

    def set(self, name, arr):
        self.state(name).upload(arr)

    # ---- stepping ----
    def TimeStep(self, n=1):
        self._chk(lib().sf_step(self.h, int(n)))

    def TimeStepLinear(self, n=1):
        self._chk(lib().sf_step_linear(self.h, int(n)))

    def Stage(self, k, linear=False):
        self._chk(lib().sf_stage(self.h, int(k), int(bool(linear))))

    def SetLinearBackground(self, u1, u2, u3, b):
        """IMEXRK::SetBackground(u1, u2, u3, b) with modal Field objects."""
        self._chk(lib().sf_set_linear_background(self.h, u1.h, u2.h if u2 is not None else None, u3.h, b.h))

    def TimeStepAdjoint(self, below, above):
        """IMEXRK::TimeStepAdjoint (IMEXRK.h:86-126).  below/above: sequences of
        modal Field objects (u1, u2, u3, b) of this solver -- the direct state at
        the end and at the start of the backward step; u2 may be None in 2D."""
        def arr(fs):
            a = (ctypes.c_void_p * 4)()
            for i, f in enumerate(fs):
                a[i] = f.h if f is not None else None
            return a
        self._chk(lib().sf_step_adjoint(self.h, arr(below), arr(above)))

    def CopyToSolver(self, u1, u2, u3, b, p=None):
        """StateVector::CopyToSolver (StateVector.h:511-525) with Field objects."""
        h = lambda f: f.h if f is not None else None
        self._chk(lib().sf_state_bind(self.h, h(u1), h(u2), h(u3), h(b), h(p)))

    def CopyFromSolver(self, u1, u2, u3, b, p=None):
        """StateVector::CopyFromSolver (StateVector.h:532-543) into Field objects."""
        h = lambda f: f.h if f is not None else None
        self._chk(lib().sf_state_fetch(self.h, h(u1), h(u2), h(u3), h(b), h(p)))

    def JoverK(self):
        """IMEXRK::JoverK (IMEXRK.h:445-459): mixing diagnostic of the current state."""
        out = ctypes.c_double()
        self._chk(lib().sf_jover_k(self.h, ctypes.byref(out)))
        return out.value

    def PrepareRunAdjoint(self):
        """The set-up of StateVector::AdjointEvolve (StateVector.cpp:158-172) with
        PrepareRunAdjoint (IMEXRK.h:183-199): FilterAll, PopulateNodalVariables,
        RemoveDivergence(0), zero pressure, PopulateNodalVariables."""
        self.FilterAll()
        self.PopulateNodalVariables()
        self.RemoveDivergence(0.0)
        self.state("p").Zero()
        self.PopulateNodalVariables()

    def RemoveDivergence(self, pressureMultiplier=1.0):
        self._chk(lib().sf_remove_divergence(self.h, float(pressureMultiplier)))

    def FilterAll(self):
        self._chk(lib().sf_filter_all(self.h))

    def PopulateNodalVariables(self):
        self._chk(lib().sf_populate_nodal(self.h))

    def energies(self):
        ke, pe = ctypes.c_double(), ctypes.c_double()
        self._chk(lib().sf_energy(self.h, ctypes.byref(ke), ctypes.byref(pe)))
        return ke.value, pe.value

    def KE(self):
        return self.energies()[0]

    def PE(self):
        return self.energies()[1]

    def InnerProd(self, a, b):
        out = ctypes.c_double()
        self._chk(lib().sf_inner_prod(a.h, b.h, ctypes.byref(out)))
        return out.value

    def RefreshNodal(self):
        """Materialise the nodal fields U1..B as the reference holds them now."""
        self._chk(lib().sf_refresh_nodal(self.h))

    def cfl_number(self, linear=False):
        """CFL (IMEXRK.h:355-374) or CFLlinear (IMEXRK.h:376-393) for the current
        deltaT, without changing it."""
        out = ctypes.c_double()
        fn = lib().sf_cfl_linear if linear else lib().sf_cfl
        self._chk(fn(self.h, ctypes.byref(out)))
        return out.value

    def CFL(self):
        """IMEXRK::CFL (IMEXRK.h:355-374): returns the CFL number for the current
        deltaT and rescales deltaT to the target CFL 0.8."""
        cfl = self.cfl_number(False)
        self.deltaT = self.deltaT * 0.8 / cfl
        return cfl

    def CFLlinear(self):
        """IMEXRK::CFLlinear (IMEXRK.h:376-393): as CFL with target 0.4, from the
        linearisation state."""
        cfl = self.cfl_number(True)
        self.deltaT = self.deltaT * 0.4 / cfl
        return cfl

    def synchronize(self):
        self._chk(lib().sf_synchronize(self.h))

    def time_steps(self, n):
        ms = ctypes.c_double()
        self._chk(lib().sf_time_steps(self.h, int(n), ctypes.byref(ms)))
        return ms.value

    def time_steps_detailed(self, n):
        """(total_ms, {class: (ms, launches)}) with one HIP event after every kernel."""
        ms = ctypes.c_double()
        k = (ctypes.c_double * len(KERNEL_CLASSES))()
        c = (ctypes.c_int * len(KERNEL_CLASSES))()
        self._chk(lib().sf_time_steps_detailed(self.h, int(n), ctypes.byref(ms), k, c))
        return ms.value, {name: (k[i], c[i]) for i, name in enumerate(KERNEL_CLASSES)}


class Batch:
    """Many independent 2D solvers advanced together (sf_batch_*).

    ``runs`` are IMEXRK objects of one precision, already set up (background,
    state, projection, dt); ``step(n)`` has the effect of ``run.TimeStep(n)``
    on every run.  The runs must stay alive while the batch exists."""

    def __init__(self, runs):
        L = lib()
        self.runs = list(runs)
        arr = (ctypes.c_void_p * len(self.runs))(*[r.h.value for r in self.runs])
        h = ctypes.c_void_p()
        rc = L.sf_batch_create(arr, len(self.runs), ctypes.byref(h))
        if rc != 0:
            raise SFError(rc, L.sf_batch_last_error(None).decode())
        self.h = h
        for r in self.runs:
            r._deps.add(self)

    def _chk(self, rc):
        if rc != 0:
            raise SFError(rc, lib().sf_batch_last_error(self.h).decode())

    def step(self, n=1):
        self._chk(lib().sf_batch_step(self.h, n))

    def time_steps(self, n):
        ms = ctypes.c_double()
        self._chk(lib().sf_batch_time_steps(self.h, n, ctypes.byref(ms)))
        return ms.value

    def info(self):
        a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        self._chk(lib().sf_batch_info(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        x, m = ctypes.c_int(), ctypes.c_int()
        t = np.zeros(6)
        self._chk(lib().sf_batch_plan(self.h, ctypes.byref(x), ctypes.byref(m), _p(t)))
        return {"launches_per_step": a.value, "xpass_groups": b.value, "modal_groups": c.value,
                "xpass_plan": XPASS_PLANS[x.value], "modal_plan": MODAL_PLANS[m.value],
                "tuned_ms_per_stage": dict(zip(["xpass_grouped", "xpass_mixed", "xpass_mixed_r8", "modal_grouped",
                                                "modal_mixed4", "modal_mixed8"], t.tolist()))}

    def select(self, xpass=-1, modal=-1):
        """Override the tuner's plan of a phase (BATCH_GROUPED, BATCH_MIXED,
        BATCH_MIXED_R8 for the x-pass, BATCH_MIXED8 for the modal phase; -1
        keeps it)."""
        self._chk(lib().sf_batch_select(self.h, int(xpass), int(modal)))

    def close(self):
        if getattr(self, "h", None):
            lib().sf_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


class Net:
    """Learned right preconditioner (sf_net_*): Conv1D layers over the vertical
    profile of every horizontal mode, BF16 on the matrix cores, trained on the GPU.

    apply(in_fields, out_fields) computes out = in + s * net(in / s), s = RMS(in);
    the field lists hold u1, u2, u3, b (u2 may be None in 2D)."""

    def __init__(self, solver, hidden=64, layers=4, seed=0):
        L = lib()
        self.solver = solver
        h = ctypes.c_void_p()
        rc = L.sf_net_create(solver.h, hidden, layers, seed, ctypes.byref(h))
        if rc != 0:
            raise SFError(rc, L.sf_net_last_error(None).decode())
        self.h = h
        solver._deps.add(self)
        nl = ctypes.c_int()
        ch = (ctypes.c_int * 16)()
        z = ctypes.c_int()
        self._chk(L.sf_net_geometry(self.h, ctypes.byref(nl), ch, ctypes.byref(z)))
        self.layers = nl.value
        self.channels = [ch[k] for k in range(nl.value + 1)]
        self.Z = z.value
        n = ctypes.c_size_t()
        self._chk(L.sf_net_param_count(self.h, ctypes.byref(n)))
        self.nparam = n.value

    def _chk(self, rc):
        if rc != 0:
            raise SFError(rc, lib().sf_net_last_error(self.h).decode())

    @staticmethod
    def _handles(fields):
        arr = (ctypes.c_void_p * 4)(*[(f.h.value if f is not None else None) for f in fields])
        return arr

    def apply(self, inp, out):
        self._chk(lib().sf_net_apply(self.h, self._handles(inp), self._handles(out)))

    def add_pair(self, w, q):
        self._chk(lib().sf_net_add_pair(self.h, self._handles(w), self._handles(q)))

    def clear_pairs(self):
        self._chk(lib().sf_net_clear_pairs(self.h))

    def samples(self):
        c = ctypes.c_int()
        self._chk(lib().sf_net_samples(self.h, ctypes.byref(c)))
        return c.value

    def train(self, epochs, batch, lr):
        a, b = ctypes.c_double(), ctypes.c_double()
        self._chk(lib().sf_net_train(self.h, epochs, batch, lr, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def get_params(self):
        p = np.zeros(self.nparam, dtype=np.float32)
        self._chk(lib().sf_net_get_params(self.h, _fp(p)))
        return p

    def set_dilations(self, dil):
        """Dilation of each layer's taps (sf_net_set_dilations)."""
        if len(dil) != self.layers:
            raise ValueError("set_dilations needs one dilation per layer (%d)" % self.layers)
        a =(ctypes.c_int * len(dil))(*[int(x) for x in dil])
        self._chk(lib().sf_net_set_dilations(self.h, a))

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.nparam
        self._chk(lib().sf_net_set_params(self.h, _fp(p)))

    def forward_raw(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        nprof = x.shape[0]
        out = np.zeros((nprof, self.Z, self.channels[-1]), dtype=np.float32)
        self._chk(lib().sf_net_forward_raw(self.h, _fp(x), nprof, _fp(out)))
        return out

    def grad_raw(self, x, target):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(target, dtype=np.float32)
        g = np.zeros(self.nparam, dtype=np.float32)
        loss = ctypes.c_double()
        self._chk(lib().sf_net_grad_raw(self.h, _fp(x), _fp(t), x.shape[0], _fp(g), ctypes.byref(loss)))
        return loss.value, g

    def close(self):
        if getattr(self, "h", None):
            lib().sf_net_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
