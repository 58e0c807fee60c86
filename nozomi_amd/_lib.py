"""ctypes binding of libnozomi_hip.so (declared in include/nozomi_hip.h).

The library is the only compute path: when it is missing or cannot be loaded
every entry point raises.  There is no CPU fallback.
"""
import ctypes
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libnozomi_hip.so")
ABI_VERSION = 5

NZ_WALK_T2 = 1
NZ_WALKER_COUNTERS = 1

_c_vp = ctypes.c_void_p


class NzSubstrateInfo(ctypes.Structure):
    """struct nz_substrate_info_t (include/nozomi_hip.h)."""
    _fields_ = [
        ("nspheres", ctypes.c_int64),
        ("nstructures", ctypes.c_int32),
        ("ncell", ctypes.c_int32 * 3),
        ("cell", ctypes.c_double * 3),
        ("margin", ctypes.c_double),
        ("ncand_total", ctypes.c_int64),
        ("ncand_max", ctypes.c_int32),
        ("ncand_mean_box", ctypes.c_double),
        ("device_bytes", ctypes.c_int64),
        ("ncand_records", ctypes.c_int64),
    ]


class NozomiError(RuntimeError):
    pass


_lib = None
_lock = threading.Lock()

# name -> (restype, argtypes); every function returns int status unless noted
_SIGS = {
    "nz_abi_version": (ctypes.c_int, []),
    "nz_last_error_string": (ctypes.c_char_p, []),
    "nz_substrate_create": (ctypes.c_int, [_c_vp, ctypes.c_int64, ctypes.c_int32, _c_vp, _c_vp,
                                           ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                           ctypes.c_double, ctypes.c_double,
                                           ctypes.POINTER(_c_vp)]),
    "nz_substrate_destroy": (ctypes.c_int, [_c_vp]),
    "nz_substrate_info": (ctypes.c_int, [_c_vp, ctypes.POINTER(NzSubstrateInfo)]),
    "nz_substrate_export_grid": (ctypes.c_int, [_c_vp, _c_vp, _c_vp, _c_vp]),
    "nz_substrate_export_reach": (ctypes.c_int, [_c_vp, _c_vp, _c_vp, _c_vp]),
    "nz_seed": (ctypes.c_int, [_c_vp, ctypes.c_int64, _c_vp, ctypes.c_int32, ctypes.c_uint64,
                               ctypes.c_int32, _c_vp, _c_vp, _c_vp]),
    "nz_walker_create": (ctypes.c_int, [_c_vp, ctypes.c_int64, ctypes.c_uint32,
                                        ctypes.POINTER(_c_vp)]),
    "nz_walker_destroy": (ctypes.c_int, [_c_vp]),
    "nz_walker_set_sort_interval": (ctypes.c_int, [_c_vp, ctypes.c_int32]),
    "nz_walker_get_step": (ctypes.c_int, [_c_vp, ctypes.POINTER(ctypes.c_int64)]),
    "nz_walker_set_step": (ctypes.c_int, [_c_vp, ctypes.c_int64]),
    "nz_walker_set_positions": (ctypes.c_int, [_c_vp, _c_vp, _c_vp]),
    "nz_walker_seed": (ctypes.c_int, [_c_vp, _c_vp, ctypes.c_int32, ctypes.c_uint64,
                                      ctypes.c_int32, _c_vp, _c_vp]),
    "nz_walk": (ctypes.c_int, [_c_vp, ctypes.c_uint64, ctypes.c_int32, ctypes.c_double, _c_vp,
                               _c_vp, _c_vp, ctypes.c_uint32, _c_vp]),
    "nz_walker_read": (ctypes.c_int, [_c_vp, _c_vp, _c_vp, _c_vp, _c_vp, _c_vp]),
    "nz_walker_read_f64": (ctypes.c_int, [_c_vp, _c_vp, _c_vp, _c_vp]),
    "nz_walker_counters": (ctypes.c_int, [_c_vp, _c_vp, _c_vp, _c_vp]),
    "nz_displacement_moments": (ctypes.c_int, [_c_vp, _c_vp, _c_vp]),
    "nz_walker_time_kernels": (ctypes.c_int, [_c_vp, ctypes.c_int32]),
    "nz_walk_record": (ctypes.c_int, [_c_vp, ctypes.c_uint64, ctypes.c_int32, ctypes.c_double, _c_vp,
                                      ctypes.c_int64, _c_vp, _c_vp]),
    "nz_walk_record_ring": (ctypes.c_int, [_c_vp, ctypes.c_uint64, ctypes.c_int32, ctypes.c_int32,
                                           ctypes.c_double, _c_vp, ctypes.c_int64, ctypes.c_int32,
                                           ctypes.c_int32, _c_vp, _c_vp]),
    "nz_waveform_block": (ctypes.c_int, [_c_vp, ctypes.c_int64, _c_vp, ctypes.c_int32, ctypes.c_int64,
                                         ctypes.c_int32, _c_vp, _c_vp, _c_vp, _c_vp]),
    "nz_phase_gemm": (ctypes.c_int, [_c_vp, ctypes.c_int64, ctypes.c_int64, _c_vp, ctypes.c_int32,
                                     _c_vp, _c_vp, _c_vp, ctypes.c_int32, ctypes.c_int32, _c_vp,
                                     _c_vp]),
    "nz_phase_init": (ctypes.c_int, [_c_vp, ctypes.c_int64, _c_vp, _c_vp, ctypes.c_int32, _c_vp,
                                     _c_vp]),
    "nz_phase_signals": (ctypes.c_int, [_c_vp, ctypes.c_int64, ctypes.c_int32, _c_vp, _c_vp]),
    "nz_batch_create": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int32, ctypes.POINTER(_c_vp)]),
    "nz_batch_destroy": (ctypes.c_int, [_c_vp]),
    "nz_batch_add": (ctypes.c_int, [_c_vp, _c_vp, ctypes.c_int64, ctypes.c_uint64, ctypes.c_int64,
                                    ctypes.c_double, _c_vp, _c_vp, ctypes.c_uint32, ctypes.c_int32,
                                    _c_vp, _c_vp, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32),
                                    _c_vp]),
    "nz_batch_remove": (ctypes.c_int, [_c_vp, ctypes.c_int32]),
    "nz_batch_run": (ctypes.c_int, [_c_vp, ctypes.c_int32, _c_vp]),
    "nz_batch_set_graph": (ctypes.c_int, [_c_vp, ctypes.c_int32]),
    "nz_batch_status": (ctypes.c_int, [_c_vp, ctypes.c_int32, ctypes.POINTER(ctypes.c_int64),
                                       ctypes.POINTER(ctypes.c_int32)]),
    "nz_batch_status_all": (ctypes.c_int, [_c_vp, _c_vp, _c_vp]),
    "nz_batch_free": (ctypes.c_int, [_c_vp, ctypes.POINTER(ctypes.c_int64),
                                     ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int32)]),
    "nz_batch_read": (ctypes.c_int, [_c_vp, ctypes.c_int32, _c_vp, _c_vp, _c_vp, _c_vp, _c_vp]),
    "nz_batch_signals": (ctypes.c_int, [_c_vp, ctypes.c_int32, ctypes.c_int32, _c_vp, _c_vp, _c_vp,
                                        _c_vp]),
    "nz_walker_kernel_time": (ctypes.c_int, [_c_vp, ctypes.POINTER(ctypes.c_double),
                                              ctypes.POINTER(ctypes.c_int64)]),
    "nz_project_phase": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int32, _c_vp, _c_vp, _c_vp,
                                        _c_vp, _c_vp]),
    "nz_signal_reduce": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int32, _c_vp, _c_vp, _c_vp,
                                        _c_vp, _c_vp]),
    "nz_signal_reduce_noalloc": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int32, _c_vp, _c_vp,
                                                _c_vp, _c_vp, _c_vp]),
}

EXPORTED = tuple(_SIGS)


def lib():
    """Load libnozomi_hip.so (raises when it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise NozomiError(
                    f"{LIB_PATH} is not built: run `python -c 'import __graft_entry__ as g; g.build()'` "
                    "(or python -m nozomi_amd.build)")
            L = ctypes.CDLL(LIB_PATH)
            for name, (res, args) in _SIGS.items():
                f = getattr(L, name)
                f.restype = res
                f.argtypes = args
            v = L.nz_abi_version()
            if v != ABI_VERSION:
                raise NozomiError(f"libnozomi_hip ABI {v}, expected {ABI_VERSION}")
            _lib = L
    return _lib


def check(rc, what=""):
    """Raise on a non-zero status: ValueError for argument errors (< 0),
    NozomiError for HIP errors (> 0)."""
    if rc == 0:
        return
    msg = lib().nz_last_error_string()
    msg = msg.decode() if msg else ""
    if rc < 0:
        raise ValueError(f"{what}: {msg}" if what else msg)
    raise NozomiError(f"{what}: HIP error {rc}: {msg}" if what else f"HIP error {rc}: {msg}")
