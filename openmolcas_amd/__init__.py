"""MI355X-native Cholesky J/K Fock build (drop-in for OpenMolcas CHO_FSCF /
CHO_FMCSCF).  The compute path is hand-written HIP for gfx950 in
``openmolcas_amd/csrc`` behind the C ABI of ``include/molcas_jk.h``."""
from ._lib import MjkError, LIB_PATH, PREC_FP64, PREC_FP32, PREC_BF16X3  # noqa: F401

__all__ = ["JKContext", "MjkError", "LIB_PATH", "PREC_FP64", "PREC_FP32", "PREC_BF16X3"]


def __getattr__(name):
    if name == "JKContext":
        from .fock import JKContext

        return JKContext
    raise AttributeError(name)
