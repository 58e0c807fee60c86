"""nz_signal_reduce_noalloc (the signal reduction that allocates nothing)
against nz_signal_reduce, called through the C ABI: two calls in a row, the
second on a smaller input (the static scratch keeps the first call's
partials), plus a host FP64 sum and the size limit of the scratch."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

_vp = ctypes.c_void_p


def _lib():
    from nozomi_amd import _lib as nl
    L = nl.lib()
    for name in ("nz_signal_reduce", "nz_signal_reduce_noalloc"):
        f = getattr(L, name)
        f.restype = ctypes.c_int
        f.argtypes = [ctypes.c_int64, ctypes.c_int32, _vp, _vp, _vp, _vp, _vp]
    return L


def _inputs(n, W, seed):
    rng = np.random.default_rng(seed)
    mom = (rng.normal(size=(3, n)) * 3.0).astype(np.float32)
    u = rng.normal(size=(3, W))
    u /= np.linalg.norm(u, axis=0)
    scale = rng.uniform(0.1, 1.0, W).astype(np.float32)
    return mom, u.astype(np.float32), scale


def _host_sums(mom, dirs, scale):
    ph = (scale[None, :] * dirs).astype(np.float32).T @ mom  # FP32 phases as the kernel forms them
    ph = ph.astype(np.float64)
    out = np.empty(2 * dirs.shape[1])
    out[0::2] = np.cos(ph).sum(axis=1)
    out[1::2] = np.sin(ph).sum(axis=1)
    return out


def test_noalloc_equals_allocating_reduction_two_calls():
    L = _lib()
    dev = torch.device("cuda:0")
    stream = _vp(torch.cuda.current_stream(dev).cuda_stream)
    for call, (n, W, seed) in enumerate(((300_001, 512, 1), (70_000, 96, 2))):
        print(f"call {call}: n={n} W={W}", flush=True)
        mom, dirs, scale = _inputs(n, W, seed)
        m = torch.from_numpy(mom).to(dev)
        d = torch.from_numpy(dirs).to(dev)
        s = torch.from_numpy(scale).to(dev)
        ref = torch.full((2 * W,), np.nan, dtype=torch.float64, device=dev)
        got = torch.full((2 * W,), np.nan, dtype=torch.float64, device=dev)
        rc = L.nz_signal_reduce(n, W, _vp(m.data_ptr()), _vp(d.data_ptr()), _vp(s.data_ptr()),
                                _vp(ref.data_ptr()), stream)
        assert rc == 0, L.nz_last_error_string()
        rc = L.nz_signal_reduce_noalloc(n, W, _vp(m.data_ptr()), _vp(d.data_ptr()),
                                        _vp(s.data_ptr()), _vp(got.data_ptr()), stream)
        assert rc == 0, L.nz_last_error_string()
        torch.cuda.synchronize(dev)
        ref_h, got_h = ref.cpu().numpy(), got.cpu().numpy()
        # same kernels, same block partition and summation order: same bits
        assert np.array_equal(ref_h, got_h)
        host = _host_sums(mom, dirs, scale)
        err = np.abs(got_h - host).max() / n
        print(f"  max |device - host FP64| / n = {err:.2e}", flush=True)
        assert err < 1e-6


def test_noalloc_refuses_more_than_its_scratch():
    L = _lib()
    dev = torch.device("cuda:0")
    m = torch.zeros((3, 16), dtype=torch.float32, device=dev)
    d = torch.zeros((3, 513), dtype=torch.float32, device=dev)
    out = torch.zeros(2 * 513, dtype=torch.float64, device=dev)
    rc = L.nz_signal_reduce_noalloc(16, 513, _vp(m.data_ptr()), _vp(d.data_ptr()), None,
                                    _vp(out.data_ptr()), _vp(0))
    assert rc < 0
    assert b"512" in L.nz_last_error_string()
