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
