"""ctypes binding of libmolcas_jk.so (the C ABI declared in include/molcas_jk.h).

The library is built in-tree by ``openmolcas_amd.build`` (hipcc, gfx950).  If
it is missing, every use raises immediately: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libmolcas_jk.so")

MJK_OK = 0
MJK_ERR_ARG = 1
MJK_ERR_HIP = 2
MJK_ERR_STATE = 3
MJK_ERR_NOMEM = 33
MJK_ERR_SHORT_READ = 77

PREC_FP64 = 0
PREC_FP32 = 1
PREC_BF16X3 = 2
PREC_LADDER = 3

PHASES = ("coulomb_v", "coulomb_j", "half_transform", "exchange", "reduce", "assembly", "pack", "backward",
          "zsave")

_c = ctypes
_vp = _c.c_void_p
_dp = _c.POINTER(_c.c_double)
_i32 = _c.c_int32
_i64 = _c.c_int64

# name -> (restype, argtypes); every entry point of include/molcas_jk.h
SIGNATURES = {
    "mjk_create": (_c.c_int, [_c.POINTER(_vp), _i32, _i64]),
    "mjk_destroy": (_c.c_int, [_vp]),
    "mjk_last_error": (_c.c_char_p, [_vp]),
    "mjk_create_shared": (_c.c_int, [_c.POINTER(_vp), _i32, _i64, _vp]),
    "mjk_device_vectors": (_c.c_int, [_vp, _c.POINTER(_vp), _c.POINTER(_i64)]),
    "mjk_set_tuning": (_c.c_int, [_vp, _i32, _i32, _i32]),
    "mjk_autotune": (_c.c_int, [_vp, _i32, _vp, _vp, _vp, _i64, _i32, _vp]),
    "mjk_set_stream": (_c.c_int, [_vp, _vp]),
    "mjk_set_x_budget": (_c.c_int, [_vp, _i64]),
    "mjk_set_reduced_set": (_c.c_int, [_vp, _i32, _i64, _vp]),
    "mjk_upload_vectors": (_c.c_int, [_vp, _i32, _i64, _i64, _vp, _i32]),
    "mjk_upload_vectors_file": (_c.c_int, [_vp, _c.c_char_p, _i32, _i64, _i64, _vp, _i32, _i64]),
    "mjk_set_ladder": (_c.c_int, [_vp, _i64, _i64]),
    "mjk_cho_fscf": (_c.c_int, [_vp, _i32, _vp, _vp, _vp, _vp, _vp, _c.c_double, _i32]),
    "mjk_cho_fscf_sym": (_c.c_int, [_vp, _i32, _vp, _i32, _vp, _vp, _vp, _vp, _vp, _c.c_double, _i32]),
    "mjk_cho_lk_scf_sym": (_c.c_int, [_vp, _i32, _vp, _i32, _vp, _vp, _vp, _vp, _vp, _vp, _c.c_double,
                                      _i32, _c.c_double, _c.c_double, _i32, _vp, _i32]),
    "mjk_cho_fock_dft_red": (_c.c_int, [_vp, _vp, _vp]),
    "mjk_cho_lk_casscf": (_c.c_int, [_vp, _vp, _vp, _vp, _vp, _c.c_double, _i32, _i32, _i32, _vp, _vp,
                                     _i32, _i32, _c.c_double, _c.c_double, _vp, _i32, _i32,
                                     _c.c_double, _i32]),
    "mjk_cho_lk_scf": (_c.c_int, [_vp, _i32, _vp, _vp, _vp, _vp, _vp, _vp, _c.c_double, _i32,
                                  _c.c_double, _c.c_double,
                                  _i32, _vp, _i32]),
    "mjk_cho_fmcscf": (_c.c_int, [_vp, _vp, _i32, _i32, _i32, _c.c_double, _vp, _i32, _vp, _i32,
                                  _vp, _vp, _vp, _vp, _i32, _i32, _c.c_double, _i32]),
    "mjk_cho_fmcscf_sym": (_c.c_int, [_vp, _i32, _vp, _vp, _vp, _vp, _vp, _c.c_double, _vp, _i32,
                                      _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _c.c_double, _i32]),
    "mjk_cho_lk_casscf_sym": (_c.c_int, [_vp, _i32, _vp, _vp, _vp, _vp, _vp, _c.c_double, _vp, _vp,
                                         _vp, _vp, _vp, _i32, _i32, _c.c_double, _c.c_double, _vp,
                                         _vp, _vp, _c.c_double, _i32]),
    "mjk_coulomb": (_c.c_int, [_vp, _i32, _vp, _i64, _vp, _i64, _vp]),
    "mjk_coulomb_noalloc": (_c.c_int, [_vp, _i32, _vp, _i64, _vp, _i64, _vp]),
    "mjk_exchange": (_c.c_int, [_vp, _i32, _vp, _vp, _vp, _i64, _i32]),
    "mjk_coulomb_exchange": (_c.c_int, [_vp, _vp, _i64, _vp, _i64, _i32, _vp, _vp, _vp, _i64, _vp]),
    "mjk_exchange_z": (_c.c_int, [_vp, _i32, _vp, _vp, _i32, _vp]),
    "mjk_coulomb_exchange_z": (_c.c_int, [_vp, _vp, _vp, _i32, _vp, _vp, _vp, _vp]),
    "mjk_exchange_mixed": (_c.c_int, [_vp, _i32, _vp, _vp, _vp, _i64, _i64, _i64]),
    "mjk_vector_norms": (_c.c_int, [_vp, _vp]),
    "mjk_exchange_screened": (_c.c_int, [_vp, _i32, _vp, _vp, _vp, _i64, _i32, _c.c_double,
                                         _c.c_double]),
    "mjk_screening_stats": (_c.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "mjk_set_lk_screening": (_c.c_int, [_vp, _i32, _c.c_double]),
    "mjk_set_shells": (_c.c_int, [_vp, _i32, _vp]),
    "mjk_set_lk_diag": (_c.c_int, [_vp, _vp, _i32, _i32, _i32]),
    "mjk_lk_stats": (_c.c_int, [_vp, _vp]),
    "mjk_choinfo_read": (_c.c_int, [_c.POINTER(_vp), _c.c_char_p, _c.c_char_p, _vp]),
    "mjk_choinfo_error": (_c.c_char_p, [_vp]),
    "mjk_choinfo_free": (_c.c_int, [_vp]),
    "mjk_choinfo_dims": (_c.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "mjk_choinfo_reduced_set": (_c.c_int, [_vp, _i32, _vp, _vp]),
    "mjk_choinfo_infvec": (_c.c_int, [_vp, _i32, _vp]),
    "mjk_load_choinfo": (_c.c_int, [_vp, _vp, _c.c_char_p, _c.c_char_p, _i64]),
    "mjk_load_choinfo_sym": (_c.c_int, [_vp, _vp, _vp, _i32, _c.c_char_p, _i64]),
    "mjk_choinfo_symmetry": (_c.c_int, [_vp, _vp, _vp, _vp]),
    "mjk_cd_incore": (_c.c_int, [_vp, _i32, _vp, _vp, _i32, _vp, _c.c_double, _vp]),
    "mjk_fock_update": (_c.c_int, [_vp, _i32, _vp, _i64, _vp, _i64, _vp, _vp, _vp, _i64, _vp]),
    "mjk_energy": (_c.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "mjk_puvx": (_c.c_int, [_vp, _i32, _vp, _vp, _i32, _i32, _vp]),
    "mjk_density": (_c.c_int, [_vp, _i32, _vp, _c.c_double, _vp]),
    "mjk_grad_orbitals": (_c.c_int, [_vp, _i32, _vp, _vp, _c.c_double, _vp]),
    "mjk_grad_vectors": (_c.c_int, [_vp, _i32, _vp, _vp, _vp, _c.c_double, _c.c_double, _vp, _i64]),
    "mjk_grad_vectors_z": (_c.c_int, [_vp, _i32, _vp, _vp, _vp, _vp, _c.c_double, _c.c_double, _vp,
                                      _i64]),
    "mjk_grad_vectors_z_mixed": (_c.c_int, [_vp, _i32, _vp, _vp, _vp, _vp, _c.c_double, _c.c_double,
                                            _vp, _i64, _i64]),
    "mjk_sync": (_c.c_int, [_vp]),
    "mjk_set_profiling": (_c.c_int, [_vp, _i32]),
    "mjk_phase_times": (_c.c_int, [_vp, _vp, _vp]),
    "mjk_measure_peak": (_c.c_int, [_vp, _i32, _vp]),
}

_lib = None


class MjkError(RuntimeError):
    def __init__(self, rc: int, msg: str):
        super().__init__(f"molcas_jk rc={rc}: {msg}")
        self.rc = rc


def load():
    """Loads the HIP library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m openmolcas_amd.build`"
                " (the J/K build has no CPU fallback)")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib
