"""libnozomi_hip.so loads and exports every entry point include/nozomi_hip.h
declares (no GPU call)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "nozomi_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(nz_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_entry_points():
    names = _declared()
    for n in ("nz_substrate_create", "nz_walk", "nz_seed", "nz_walker_create", "nz_signal_reduce"):
        assert n in names


def test_library_exports_every_declared_symbol():
    from nozomi_amd import _lib
    L = _lib.lib()
    for name in _declared():
        assert hasattr(L, name), name
    assert L.nz_abi_version() == _lib.ABI_VERSION
    assert set(_lib.EXPORTED) == set(_declared())


def test_argument_errors_are_reported():
    from nozomi_amd import _lib
    L = _lib.lib()
    h = ctypes.c_void_p()
    rc = L.nz_substrate_create(None, 0, 0, None, None, 10.0, 10.0, 10.0, 0.0, 0.0, ctypes.byref(h))
    assert rc < 0
    assert b"structure" in L.nz_last_error_string()
    rc = L.nz_walk(None, 0, 1, 0.01, None, None, None, 0, None)
    assert rc < 0 and b"walker" in L.nz_last_error_string()
    w = ctypes.c_void_p()
    rc = L.nz_walker_create(None, 10, 0, ctypes.byref(w))
    assert rc < 0 and b"substrate" in L.nz_last_error_string()
    b = ctypes.c_void_p()
    rc = L.nz_batch_create(0, 4, ctypes.byref(b))
    assert rc < 0 and b"capacity" in L.nz_last_error_string()
    rc = L.nz_batch_run(None, 2, None)
    assert rc < 0 and b"batch" in L.nz_last_error_string()
    sim = ctypes.c_int32()
    rc = L.nz_batch_add(None, None, 10, 0, 10, 0.01, None, None, 0, 64, None, None, 0,
                        ctypes.byref(sim), None)
    assert rc < 0 and b"batch" in L.nz_last_error_string()
    # the MFMA contraction's shape rules are checked before any launch (the
    # buffers below are host memory and are never touched)
    buf = ctypes.create_string_buffer(64)
    p = ctypes.cast(buf, ctypes.c_void_p)
    rc = L.nz_phase_gemm(p, 192, 128, p, 64, p, p, None, 50, 0, p, None)
    assert rc < 0 and b"multiples of 4" in L.nz_last_error_string()
    rc = L.nz_phase_gemm(p, 240, 128, p, 80, p, p, None, 64, 0, p, None)
    assert rc < 0 and b"K = 32, 64, 96 or 128" in L.nz_last_error_string()


def test_host_tensors_are_refused_before_any_launch():
    """A host (CPU) tensor, or a strided view, handed where a kernel expects
    device memory raises ValueError in the Python wrapper: the library is
    never called with a host address (that would fault the GPU)."""
    import pytest
    import torch
    from nozomi_amd import runtime as rt
    with pytest.raises(ValueError, match="GPU tensor"):
        rt._ptr(torch.zeros(6, dtype=torch.float64), None, "stepmom")
    with pytest.raises(ValueError, match="torch.Tensor"):
        rt._ptr([0.0] * 6, None, "stepmom")
    mom = torch.zeros((3, 16), dtype=torch.float32)
    dirs = torch.zeros((3, 4), dtype=torch.float32)
    with pytest.raises(ValueError, match="GPU tensor"):
        rt.signal_sums(mom, dirs)
    with pytest.raises(ValueError, match="GPU tensor"):
        rt.project_phase(mom, dirs)
    with pytest.raises(ValueError, match="shape"):
        rt.signal_sums(mom.T, dirs)
