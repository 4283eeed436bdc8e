"""ctypes binding of libedlat_hip.so (the C-ABI declared in include/edlat_gpu.h).

The shared library is built in-tree by ``__graft_entry__.build()`` (or ``make``
in this directory).  Loading fails loudly when it is missing: there is no CPU
fallback on the product path.
"""
import ctypes
import os
from ctypes import POINTER, c_char_p, c_double, c_int, c_int32, c_int64, c_uint64, c_void_p

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libedlat_hip.so")

EDL_FP64, EDL_FP32, EDL_MIXED = 0, 1, 2
EDL_OP_C, EDL_OP_CDG, EDL_OP_SZ, EDL_OP_J = 1, 2, 3, 4


class SeedDesc(ctypes.Structure):
    _fields_ = [("op", c_int), ("ispin", c_int), ("io", c_int), ("jo", c_int)]


_dp = POINTER(c_double)
_ip = POINTER(c_int)
_i32p = POINTER(c_int32)
_i64p = POINTER(c_int64)

# name -> (restype, argtypes); every entry point of include/edlat_gpu.h
SIGNATURES = {
    "edl_ctx_create": (c_int, [POINTER(c_void_p), c_int]),
    "edl_ctx_destroy": (c_int, [c_void_p]),
    "edl_last_error": (c_char_p, [c_void_p]),
    "edl_set_hamiltonian": (c_int, [c_void_p, c_int, c_int, c_int, _ip, _dp, _dp, _dp, c_double, c_double,
                                    c_double, c_double, c_double, c_int, c_int]),
    "edl_sector_index": (c_int, [c_int, c_int, c_int]),
    "edl_sector_qn": (c_int, [c_int, c_int, _ip, _ip]),
    "edl_sector_build": (c_int, [c_void_p, c_int, _i64p, _i32p, _i32p]),
    "edl_sector_maps": (c_int, [c_void_p, c_int, _i32p, _i32p]),
    "edl_sector_hops": (c_int, [c_void_p, c_int, c_int, _i32p, _i32p, _dp]),
    "edl_sector_rank": (c_int, [c_void_p, c_int, c_int, c_int64, _i32p, _i32p]),
    "edl_hxv": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p]),
    "edl_hxv_noalloc": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p]),
    "edl_hxv_host": (c_int, [c_void_p, c_int, c_int, _dp, _dp]),
    "edl_hxv_host_complex": (c_int, [c_void_p, c_int, c_int, _dp, _dp]),
    "edl_vec_random": (c_int, [c_void_p, c_int, c_void_p, c_int64, c_uint64]),
    "edl_vec_dot": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_int64, _dp]),
    "edl_device_alloc": (c_int, [c_void_p, c_int64, POINTER(c_void_p)]),
    "edl_copy_h2d": (c_int, [c_void_p, c_void_p, c_void_p, c_int64]),
    "edl_copy_d2h": (c_int, [c_void_p, c_void_p, c_void_p, c_int64]),
    "edl_lanczos_gs": (c_int, [c_void_p, c_int, c_int, c_int, c_double, c_uint64, c_double, _dp, _ip, _dp, _dp,
                               _dp]),
    "edl_lanczos_states": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_double, c_uint64, _dp, _dp]),
    "edl_lanczos_stats": (c_int, [c_void_p, _ip, _ip, _dp]),
    "edl_set_state": (c_int, [c_void_p, c_int, c_double, _dp]),
    "edl_get_state": (c_int, [c_void_p, _dp]),
    "edl_tridiag": (c_int, [c_void_p, c_int, c_int, _dp, c_int, _dp, _dp, _dp, _ip]),
    "edl_tridiag_batched": (c_int, [c_void_p, c_int, POINTER(SeedDesc), c_int, c_int, c_int, _dp, _dp, _dp, _ip,
                                    _ip]),
    "edl_lanczos_poles": (c_int, [c_int, _dp, _dp, c_double, c_double, c_int, _dp, _dp]),
    "edl_observables": (c_int, [c_void_p, _dp, _dp, _dp]),
    "edl_local_energy": (c_int, [c_void_p, _dp]),
    "edl_correlations": (c_int, [c_void_p, _dp, _dp]),
    "edl_last_timing": (c_int, [c_void_p, _dp, _i64p]),
    "edl_bench_hxv": (c_int, [c_void_p, c_int, c_int, c_int, _dp]),
    "edl_bench_passes": (c_int, [c_void_p, c_int, c_int, c_int, _dp, _ip]),
}

_lib = None


def load():
    """Load libedlat_hip.so once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with __graft_entry__.build() (no CPU fallback exists)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib
