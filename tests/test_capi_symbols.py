"""The C ABI library loads and exports every entry point include/molcas_jk.h
declares (CPU only: no device calls)."""
import ctypes
import os
import re

import pytest

from openmolcas_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "molcas_jk.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mjk_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_expected_surface():
    names = header_functions()
    for must in ("mjk_create", "mjk_destroy", "mjk_set_reduced_set", "mjk_upload_vectors",
                 "mjk_cho_fscf", "mjk_cho_fmcscf", "mjk_coulomb", "mjk_exchange",
                 "mjk_fock_update", "mjk_energy", "mjk_puvx"):
        assert must in names


def test_binding_covers_header_exactly():
    assert sorted(_lib.SIGNATURES) == header_functions()


@pytest.mark.skipif(not os.path.exists(_lib.LIB_PATH), reason="library not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in header_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_library_is_built():
    # build() runs before the CPU suite; a missing library is a build failure
    assert os.path.exists(_lib.LIB_PATH), "build the library first: python -m openmolcas_amd.build"
