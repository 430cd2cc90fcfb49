"""ctypes binding of libfdtdpic.so (include/fdtdpic.h).

The library is built in-tree by `__graft_entry__.build()` (or
`python -m fdtd_pic_amd.build`) into this directory.  There is no fallback:
if the library is missing or no GPU is visible, calls raise FdtdError.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfdtdpic.so")

FDTD_OK = 0
FDTD_E_ARG, FDTD_E_HIP, FDTD_E_STATE, FDTD_E_NOMEM = -1, -2, -3, -4
F64, F32 = 0, 1
SLAB_NONE, SLAB_DIEL, SLAB_LINEAR, SLAB_KERR_REF, SLAB_KERR_E = 0, 1, 2, 3, 4
MAX_PROBES = 16
FIELD = dict(E=0, H=1, D=2, P=3, PM=4, PSIE=5, PSIH=6, J=7)

_dp = C.POINTER(C.c_double)


class FdtdError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"libfdtdpic error {code}: {msg}")
        self.code = code


class SimDesc(C.Structure):
    _fields_ = [
        ("N", C.c_int32), ("steps", C.c_int32),
        ("ke0", C.c_double), ("kh0", C.c_double),
        ("Ls", C.c_int32), ("Rs", C.c_int32),
        ("s_ke", _dp), ("s_kh", _dp), ("s_bE", _dp), ("s_cE", _dp),
        ("s_CbE", _dp), ("s_bH", _dp), ("s_cH", _dp), ("s_C2H", _dp),
        ("F", C.c_int32), ("R", C.c_int32), ("slab_kind", C.c_int32), ("ade", C.c_int32),
        ("keS", C.c_double), ("khS", C.c_double), ("kD", C.c_double), ("jD", C.c_double),
        ("sl_A", _dp), ("sl_B", _dp), ("sl_C", _dp), ("sl_e", _dp),
        ("kerr_a", C.c_double), ("kerr_b", C.c_double), ("kerr_c", C.c_double),
        ("kerr_e0", C.c_double), ("kerr_lin", C.c_double), ("kerr_chi3", C.c_double),
        ("src_cell", C.c_int32), ("tfsf", C.c_int32),
        ("srcE", _dp), ("srcH", _dp),
        ("nprobe", C.c_int32),
        ("probe_cell", C.c_int32 * MAX_PROBES),
        ("probe_t0", C.c_int32 * MAX_PROBES), ("probe_t1", C.c_int32 * MAX_PROBES),
        ("snap_interval", C.c_int32), ("snap_rows", C.c_int32),
        ("has_current", C.c_int32),
    ]


class BeamDesc(C.Structure):
    _fields_ = [
        ("n", C.c_int64),
        ("q", C.c_double), ("m", C.c_double), ("w", C.c_double),
        ("dt", C.c_double), ("dz", C.c_double),
        ("x", _dp), ("u", _dp),
        ("group", C.c_int32),
    ]


# every symbol include/fdtdpic.h declares, with (restype, argtypes)
_vp = C.c_void_p
SIGNATURES = {
    "fdtd_ctx_create": (C.c_int, [C.c_int, C.POINTER(_vp)]),
    "fdtd_ctx_destroy": (C.c_int, [_vp]),
    "fdtd_set_tiling": (C.c_int, [_vp, C.c_int, C.c_int]),
    "fdtd_set_tiling_mixed": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int]),
    "fdtd_set_pic_fusion": (C.c_int, [_vp, C.c_int]),
    "fdtd_set_pic_lds_rows": (C.c_int, [_vp, C.c_int]),
    "fdtd_set_persistent": (C.c_int, [_vp, C.c_int]),
    "fdtd_persistent_active": (C.c_int, [_vp, C.POINTER(C.c_int)]),
    "fdtd_set_halo_exchange": (C.c_int, [_vp, C.c_int]),
    "fdtd_halo_exchange_active": (C.c_int, [_vp, C.POINTER(C.c_int)]),
    "fdtd_add_sim": (C.c_int, [_vp, C.POINTER(SimDesc), C.POINTER(C.c_int)]),
    "fdtd_remove_sim": (C.c_int, [_vp, C.c_int]),
    "fdtd_sim_progress": (C.c_int, [_vp, C.c_int, C.POINTER(C.c_int64)]),
    "fdtd_reset_sim": (C.c_int, [_vp, C.c_int]),
    "fdtd_state_size": (C.c_int, [_vp, C.c_int, C.POINTER(C.c_int64)]),
    "fdtd_state_save": (C.c_int, [_vp, C.c_int, _vp, C.c_int64]),
    "fdtd_state_load": (C.c_int, [_vp, C.c_int, _vp, C.c_int64]),
    "fdtd_run": (C.c_int, [_vp, C.c_int64]),
    "fdtd_run_all": (C.c_int, [_vp]),
    "fdtd_sync": (C.c_int, [_vp]),
    "fdtd_busy": (C.c_int, [_vp, C.POINTER(C.c_int)]),
    "fdtd_read_field": (C.c_int, [_vp, C.c_int, C.c_int, _dp, C.c_int64]),
    "fdtd_write_field": (C.c_int, [_vp, C.c_int, C.c_int, _dp, C.c_int64]),
    "fdtd_read_probes": (C.c_int, [_vp, C.c_int, _dp, C.c_int64]),
    "fdtd_read_snapshots": (C.c_int, [_vp, C.c_int, _dp, C.c_int64]),
    "fdtd_last_error": (C.c_char_p, []),
    "fdtd_timing": (C.c_int, [_vp, C.POINTER(C.c_double), C.POINTER(C.c_int64), C.c_int]),
    "fdtd_set_profile": (C.c_int, [_vp, C.c_int]),
    "fdtd_device_bytes": (C.c_int, [_vp, C.POINTER(C.c_int64)]),
    "fdtd_version": (C.c_char_p, []),
    "fdtd_kernel_timing": (C.c_int, [_vp, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
    "fdtd_set_energy": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int]),
    "fdtd_read_energy": (C.c_int, [_vp, C.c_int, _dp, C.POINTER(C.c_int64), C.c_int64, C.POINTER(C.c_int64)]),
    "fdtd_probe_spectra": (C.c_int, [_vp, C.c_int, C.POINTER(C.c_int), C.c_int, _dp, C.c_int, C.c_double, _dp]),
    "fdtd_probe_spectra_noalloc": (C.c_int, [_vp, C.c_int, C.POINTER(C.c_int), C.c_int, _dp, C.c_int, C.c_double,
                                             _dp]),
    "fdtd_step_op": (C.c_int, [_vp, C.c_int, C.c_int, C.POINTER(_dp), C.POINTER(C.c_int64), C.c_int, _dp,
                              C.POINTER(C.c_int32)]),
    "pic_attach": (C.c_int, [_vp, C.c_int, C.POINTER(BeamDesc)]),
    "pic_read": (C.c_int, [_vp, C.c_int, _dp, _dp, C.POINTER(C.c_int8), C.c_int64]),
}

_lib = None

def load(path=None):
    """Load libfdtdpic.so and bind every declared symbol (raises if missing)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise FdtdError(FDTD_E_STATE, f"{p} is not built; run __graft_entry__.build()")
    lib = C.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def check(rc):
    if rc != FDTD_OK:
        raise FdtdError(rc, load().fdtd_last_error().decode())
    return rc


def dptr(a):
    return a.ctypes.data_as(_dp)


def as_f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)
