"""Builds libmolcas_jk.so in-tree with hipcc for gfx950.

Usage: ``python -m openmolcas_amd.build``.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "csrc", "_build")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libmolcas_jk.so")

SOURCES = ["capi.hip", "coulomb.hip", "exchange.hip", "assembly.hip", "puvx.hip", "backward.hip", "peaks.hip", "tune.hip",
           "deco.hip", "lk.hip", "chofiles.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC",
    # packed FP32 instructions are kept out of every code object
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # MFMA accumulators in ordinary VGPRs: otherwise the compiler keeps a
    # loop-carried accumulator in VGPRs but issues the MFMAs on AGPR copies,
    # moving every accumulator across at each trip and draining the MFMA
    # pipeline before the copy back
    "-mllvm", "-amdgpu-mfma-vgpr-form=1",
    "-Wall", "-Wno-unused-result", "-Wno-unused-value", "-Wno-pass-failed",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
]


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed for {src}")
    return obj


FORTRAN = os.path.join(HERE, "fortran")
FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")
FORTRAN_BIN = os.path.join(FORTRAN, "bin", "fscf_driver")
FORTRAN_SYM_BIN = os.path.join(FORTRAN, "bin", "fscf_sym_driver")
FORTRAN_UHF_CAS_BIN = os.path.join(FORTRAN, "bin", "uhf_cas_driver")
FORTRAN_CAS_SYM_BIN = os.path.join(FORTRAN, "bin", "fmcscf_sym_driver")
FORTRAN_CHOINFO_BIN = os.path.join(FORTRAN, "bin", "choinfo_driver")


def build_fortran(verbose: bool = False):
    """ISO_C_BINDING interface module + the Fortran drivers of the drop-ins:
    CHO_FSCF C1 and with symmetry, the two-density UHF / CASSCF calls on
    separately allocated arrays, CHO_FMCSCF / CHO_LK_CASSCF with symmetry, and
    a caller that takes its vectors from the reference's Cholesky files
    (amdflang).  Returns the C1 driver path, or None without a compiler."""
    if not os.path.exists(FLANG):
        if verbose:
            print("amdflang not found: Fortran interface not built")
        return None
    obj = os.path.join(FORTRAN, "_build")
    os.makedirs(obj, exist_ok=True)
    os.makedirs(os.path.dirname(FORTRAN_BIN), exist_ok=True)
    iface_o = os.path.join(obj, "molcas_jk_iface.o")
    cmds = [
        [FLANG, "-O2", "-J", obj, "-c", os.path.join(FORTRAN, "molcas_jk_iface.F90"), "-o", iface_o],
        [FLANG, "-O2", "-J", obj, os.path.join(FORTRAN, "fscf_driver.F90"), iface_o, "-L" + LIBDIR,
         "-lmolcas_jk", "-Wl,-rpath,$ORIGIN/../../lib", "-o", FORTRAN_BIN],
        [FLANG, "-O2", "-J", obj, os.path.join(FORTRAN, "fscf_sym_driver.F90"), iface_o, "-L" + LIBDIR,
         "-lmolcas_jk", "-Wl,-rpath,$ORIGIN/../../lib", "-o", FORTRAN_SYM_BIN],
        [FLANG, "-O2", "-J", obj, os.path.join(FORTRAN, "uhf_cas_driver.F90"), iface_o, "-L" + LIBDIR,
         "-lmolcas_jk", "-Wl,-rpath,$ORIGIN/../../lib", "-o", FORTRAN_UHF_CAS_BIN],
        [FLANG, "-O2", "-J", obj, os.path.join(FORTRAN, "fmcscf_sym_driver.F90"), iface_o, "-L" + LIBDIR,
         "-lmolcas_jk", "-Wl,-rpath,$ORIGIN/../../lib", "-o", FORTRAN_CAS_SYM_BIN],
        [FLANG, "-O2", "-J", obj, os.path.join(FORTRAN, "choinfo_driver.F90"), iface_o, "-L" + LIBDIR,
         "-lmolcas_jk", "-Wl,-rpath,$ORIGIN/../../lib", "-o", FORTRAN_CHOINFO_BIN],
    ]
    for cmd in cmds:
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError("amdflang failed")
    if verbose:
        print("built", FORTRAN_BIN)
    return FORTRAN_BIN


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc link failed")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    build_fortran(verbose)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
