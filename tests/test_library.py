"""CPU-side checks of the product library: it loads and exports the C ABI."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "sfhip.h")).read()
    return sorted(set(re.findall(r"^(?:int|const char\*)\s+(sf_\w+)\(", src, re.M)))


def test_header_matches_wrapper():
    import stratiflow_amd as sf
    assert sorted(sf.EXPORTED) == _declared()


def test_library_exports_every_symbol():
    import stratiflow_amd as sf
    if not os.path.exists(sf.LIB_PATH):
        pytest.skip("libsfhip.so not built")
    L = ctypes.CDLL(sf.LIB_PATH)
    for name in _declared():
        assert hasattr(L, name), name


def test_create_without_gpu_fails_cleanly():
    """Without a GPU the context cannot be created; the error is a status code, not a crash."""
    import stratiflow_amd as sf
    if not os.path.exists(sf.LIB_PATH):
        pytest.skip("libsfhip.so not built")
    try:
        import torch  # noqa: F401
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if has_gpu:
        pytest.skip("GPU present")
    with pytest.raises(sf.SFError):
        sf.IMEXRK(48, 1, 64, sf.TWO_D, 14.28, 1.0, 10.0, 1000.0, 0.1, 1.0)


def test_null_handles_return_argument_errors():
    """Every entry point that takes a handle answers NULL with SF_ERR_ARG (or the
    documented status) before touching the device, so this runs without a GPU."""
    import stratiflow_amd as sf
    if not os.path.exists(sf.LIB_PATH):
        pytest.skip("libsfhip.so not built")
    L = ctypes.CDLL(sf.LIB_PATH)
    vp = ctypes.c_void_p
    d = ctypes.c_double()
    i = ctypes.c_int()
    buf = (ctypes.c_double * 8)()
    assert L.sf_field_free(vp()) == 0  # freeing NULL is a no-op, like free(NULL)
    calls = {
        "sf_field_zero": (None,),
        "sf_field_scale": (None, ctypes.c_double(2.0)),
        "sf_field_upload": (None, buf),
        "sf_field_download": (None, buf),
        "sf_field_copy": (None, None),
        "sf_to_modal": (None, None),
        "sf_to_nodal": (None, None),
        "sf_ddx": (None, None),
        "sf_ddz": (None, None),
        "sf_inner_prod": (None, None, ctypes.byref(d)),
        "sf_dot_f64": (None, None, ctypes.byref(d)),
        "sf_field_enforce_bcs": (None,),
        "sf_phase_shift": (None, ctypes.c_double(0.5)),
        "sf_remove_average": (None,),
        "sf_field_add_profile": (None, buf),
        "sf_field_mean_profile": (None, buf),
        "sf_step": (None, 1),
        "sf_step_linear": (None, 1),
        "sf_step_adjoint": (None, None, None),
        "sf_jover_k": (None, ctypes.byref(d)),
        "sf_state_bind": (None, None, None, None, None, None),
        "sf_state_fetch": (None, None, None, None, None, None),
        "sf_set_dt": (None, ctypes.c_double(0.01)),
        "sf_energy": (None, ctypes.byref(d), ctypes.byref(d)),
        "sf_cfl": (None, ctypes.byref(d)),
        "sf_net_set_dilations": (None, (ctypes.c_int * 4)(1, 2, 4, 8)),
        "sf_net_samples": (None, ctypes.byref(i)),
        "sf_batch_step": (None, 1),
    }
    bad = {}
    for name, args in calls.items():
        fn = getattr(L, name)
        fn.restype = ctypes.c_int
        rc = fn(*[a if a is not None else vp() for a in args])
        if rc != 1:  # SF_ERR_ARG
            bad[name] = rc
    assert not bad, bad
