"""The CHO_FSCF drop-in called from Fortran through the ISO_C_BINDING module
(openmolcas_amd/fortran), C1 and with two irreps, checked against the oracle
(FP64, relative 1e-11)."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from openmolcas_amd.build import FORTRAN_BIN, FORTRAN_SYM_BIN
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


@pytest.mark.skipif(not os.path.exists(FORTRAN_SYM_BIN), reason="Fortran driver not built (no amdflang)")
def test_fortran_cho_fscf_symmetry_dropin(gpu, tmp_path):
    from _sym import NBAS, make_problem
    p = make_problem(1, 0.5, seed=31)
    rs1, rs2 = (np.array(r, dtype=np.float64) for r in p["rs"])
    nocc = [p["nforb"][0][s] + p["niorb"][0][s] for s in range(2)]
    # Porb(1)%SB(s)%A2(nOcc(s),nBas(s)) blocks: (nbas, nocc) C-order = orbital index fastest
    porb = np.concatenate([np.ascontiguousarray(b).reshape(-1) for b in p["porb"][0]])
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(fin, "wb") as f:
        np.array(NBAS + [p["V1"].shape[0], p["V2"].shape[0], rs1.size, rs2.size] + nocc + [0.5],
                 dtype=np.float64).tofile(f)
        for a in (rs1, rs2, p["V1"], p["V2"], porb, p["dsym"], p["fsym"][0]):
            np.ascontiguousarray(a, dtype=np.float64).tofile(f)
    r = subprocess.run([FORTRAN_SYM_BIN, str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    flt = np.fromfile(fout, dtype=np.float64)
    want = p["want"][0]
    assert np.max(np.abs(flt - want)) / np.max(np.abs(want)) < 1e-11
