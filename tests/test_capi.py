"""The C-ABI library loads and exports every entry point include/mpmx.h declares
(no compute calls: runs without a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "mpmx.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(mpmx_\w+)\s*\(", src, re.M)))


def test_header_declares_entry_points():
    names = _declared()
    assert "mpmx_step" in names and "mpmx_create" in names
    from mpm_amd import capi
    assert sorted(capi.EXPORTS) == names


def test_library_exports_every_declared_symbol():
    from mpm_amd import capi
    if not os.path.exists(capi.LIB_PATH):
        pytest.skip("libmpmx.so not built")
    lib = ctypes.CDLL(capi.LIB_PATH)
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_create_without_device_fails_cleanly():
    """On a host without a GPU, creating a solver reports MPMX_E_HIP."""
    import torch  # noqa: F401  (only to detect the device the same way the runtime does)
    from mpm_amd import capi
    if not os.path.exists(capi.LIB_PATH):
        pytest.skip("libmpmx.so not built")
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(capi.MpmxError) as e:
        capi.Solver(2)
    assert e.value.code == capi.E_HIP


def _declared_pod():
    src = open(os.path.join(ROOT, "include", "mpmx_pod.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(mpmx_pod_\w+)\s*\(", src, re.M)))


def test_checkpoint_library_exports_every_declared_symbol():
    """include/mpmx_pod.h is served by mpm_amd/libmpmx_pod.so (host only)."""
    from mpm_amd import pod
    names = _declared_pod()
    assert {"mpmx_pod_write", "mpmx_pod_read", "mpmx_pod_count"} <= set(names)
    if not os.path.exists(pod.LIB_PATH):
        pytest.skip("libmpmx_pod.so not built")
    lib = ctypes.CDLL(pod.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
