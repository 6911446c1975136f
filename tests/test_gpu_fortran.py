"""The CHO_FSCF drop-in called from Fortran through the ISO_C_BINDING module
(openmolcas_amd/fortran), checked against the oracle (FP64, relative 1e-11)."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from openmolcas_amd.build import FORTRAN_BIN
from openmolcas_amd.synthetic import chol_vectors_np, fold, npair, orbitals_np

pytestmark = pytest.mark.gpu


@pytest.mark.skipif(not os.path.exists(FORTRAN_BIN), reason="Fortran driver not built (no amdflang)")
def test_fortran_cho_fscf_dropin(gpu, tmp_path):
    n, M, o = 41, 33, 6
    L = chol_vectors_np(n, M, seed=90)
    C = orbitals_np(n, o, seed=91)
    porb = np.sqrt(2.0) * C                   # memory of Porb A2(nOcc, nBas)
    dlt = fold(2.0 * C @ C.T)
    flt0 = np.random.default_rng(92).standard_normal(npair(n))
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(fin, "wb") as f:
        np.array([n, M, o, 0.5], dtype=np.float64).tofile(f)
        for a in (L, porb, dlt, flt0):
            np.ascontiguousarray(a, dtype=np.float64).tofile(f)
    r = subprocess.run([FORTRAN_BIN, str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    flt = np.fromfile(fout, dtype=np.float64)
    want = oracle.cho_fscf(oracle.CholStore.full(L), flt0[None], [o], [porb], dlt, 0.5)[0]
    assert np.max(np.abs(flt - want)) / np.max(np.abs(want)) < 1e-11
