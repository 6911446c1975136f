"""The drop-ins called from Fortran through the ISO_C_BINDING module
(openmolcas_amd/fortran): CHO_FSCF C1 and with two irreps, and the
two-density calls (UHF, CASSCF FI/FA) on separately allocated arrays and
CHO_FMCSCF with two irreps, checked against the oracle (FP64, relative
1e-11); and a Fortran caller that reads the reference's Cholesky files
(CHORST, CHRED, CHVEC1..4 of four irreps) and runs CHO_FSCF, bit-identical
to the Python caller on the same files."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from openmolcas_amd.build import (FORTRAN_BIN, FORTRAN_CAS_SYM_BIN, FORTRAN_CHOINFO_BIN, FORTRAN_SYM_BIN,
                                  FORTRAN_UHF_CAS_BIN)
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


@pytest.mark.skipif(not os.path.exists(FORTRAN_UHF_CAS_BIN), reason="Fortran driver not built (no amdflang)")
def test_fortran_two_density_separate_arrays(gpu, tmp_path):
    """UHF (nDen = 2) CHO_FSCF / CHO_LK_SCF and CHO_FMCSCF / CHO_LK_CASSCF with
    FI/FA called from Fortran with every per-density array allocated on its
    own and followed by canary words (src/scf/choscf_drv.F90:420,433,
    src/fock_util/cho_cas_drv.F90:277,285): results against the oracle and
    no canary touched."""
    n, M = 37, 29
    oa, ob, ni, na = 6, 4, 5, 4
    nfro, norb = 1, n - 1
    exfac, factxi = 0.8, -0.5
    L = chol_vectors_np(n, M, seed=93)
    rng = np.random.default_rng(94)
    Ca, Cb = orbitals_np(n, oa, seed=95), orbitals_np(n, ob, seed=96)
    dlt = fold(Ca @ Ca.T + Cb @ Cb.T)
    flta0, fltb0 = rng.standard_normal(npair(n)), rng.standard_normal(npair(n))
    cmo, _ = np.linalg.qr(rng.standard_normal((n, n)))
    Ci, Cact = cmo[:, :ni], cmo[:, ni:ni + na]
    occ = rng.uniform(0.1, 1.9, na)
    DI, DA = fold(2.0 * Ci @ Ci.T), fold((Cact * occ) @ Cact.T)
    fi0, fa0 = rng.standard_normal(npair(n)), rng.standard_normal(npair(n))
    p1, p2 = np.sqrt(2.0) * Ci, Cact * np.sqrt(occ)

    def square(c):  # A2(nBas,nBas), AO index fastest
        sq = np.zeros((n, n))
        sq[:, :c.shape[1]] = c
        return np.ascontiguousarray(sq.T)

    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(fin, "wb") as f:
        np.array([n, M, oa, ob, ni, na, na, nfro, norb, exfac, factxi], dtype=np.float64).tofile(f)
        for a in (L, Ca, Cb, square(Ca), square(Cb), dlt, flta0, fltb0, DI, DA, fi0, fa0, p1, p2, Cact,
                  cmo.T, square(p1), square(p2)):
            np.ascontiguousarray(a, dtype=np.float64).tofile(f)
    r = subprocess.run([FORTRAN_UHF_CAS_BIN, str(fin), str(fout)], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    out = np.fromfile(fout, dtype=np.float64)
    npr, nw = npair(n), norb * na * (na * (na + 1) // 2)
    sizes = [npr] * 8 + [nw] + [npr] * 2 + [nw]
    parts = np.split(out[:-1], np.cumsum(sizes)[:-1])
    assert out[-1] == 0.0, f"{int(out[-1])} canary words after the per-density arrays were written"
    st = oracle.CholStore.full(L)
    want = oracle.cho_fscf(st, np.stack([flta0, fltb0]), [oa, ob], [Ca, Cb], dlt, exfac)
    lk_f, lk_k = oracle.cho_lk_scf(st, np.stack([flta0, fltb0]), [oa, ob], [Ca, Cb], dlt, exfac)
    fm_f, fm_w = oracle.cho_fmcscf(st, np.stack([fi0, fa0]), ni, na, factxi, np.stack([DI, DA]), True,
                                   p1, na, p2, Cact, cmo, nfro, norb)
    lc_f, lc_w = oracle.cho_lk_casscf(st, np.stack([DI, DA]), np.stack([fi0, fa0]), p1, factxi, ni, na, na,
                                      Cact, p2, True, cmo, nfro, norb, 1.0)
    expect = [want[0], want[1], lk_f[0], lk_f[1], lk_k[0], lk_k[1], fm_f[0], fm_f[1], fm_w,
              lc_f[0], lc_f[1], lc_w]
    for got, exp in zip(parts, expect):
        assert np.max(np.abs(got - exp)) / max(np.max(np.abs(exp)), 1e-300) < 1e-11


@pytest.mark.skipif(not os.path.exists(FORTRAN_CAS_SYM_BIN), reason="Fortran driver not built (no amdflang)")
@pytest.mark.parametrize("do_active", [True, False])
def test_fortran_cho_fmcscf_symmetry_dropin(gpu, tmp_path, do_active):
    """CHO_FMCSCF and unscreened CHO_LK_CASSCF with nSym = 2 from Fortran
    (FI/FA, DI/DA separate arrays): FI, FA and W_PWXY of each against the C1
    oracle and the dense (pw|xy)."""
    from _sym import NBAS, make_casscf_problem
    q = make_casscf_problem(do_active, seed=41)
    p = q["vectors"]
    rs1, rs2 = (np.array(r, dtype=np.float64) for r in p["rs"])
    nfi = [q["nforb"][s] + q["niorb"][s] for s in range(2)]

    def blocks(mats, ks):  # A2(nk(s), nBas(s)) irrep blocks back to back
        return np.concatenate([np.ascontiguousarray(mats[s][:, :ks[s]]).reshape(-1) for s in range(2)])

    def squares(mats, ks):  # A2(nBas(s), nBas(s)) blocks, AO index fastest, first ks(s) columns
        out = []
        for s in range(2):
            sq = np.zeros((NBAS[s], NBAS[s]))
            sq[:, :ks[s]] = mats[s][:, :ks[s]]
            out.append(np.ascontiguousarray(sq.T).reshape(-1))
        return np.concatenate(out)

    nw = q["want_w"].size
    hdr = (NBAS + [p["V1"].shape[0], p["V2"].shape[0], rs1.size, rs2.size] + q["nforb"] + q["niorb"]
           + q["naorb"] + q["nchm"] + q["nfro"] + q["norb"] + [q["factxi"], float(do_active), nw])
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(fin, "wb") as f:
        np.array(hdr, dtype=np.float64).tofile(f)
        for a in (rs1, rs2, p["V1"], p["V2"], q["dsym"][0], q["dsym"][1], q["fsym0"][0], q["fsym0"][1],
                  blocks(q["porb1"], nfi), blocks(q["porb2"], q["nchm"]), blocks(q["porb3"], q["naorb"]),
                  np.concatenate([np.ascontiguousarray(c.T).reshape(-1) for c in q["cmo"]]),
                  squares(q["porb1"], nfi), squares(q["porb2"], q["nchm"])):
            np.ascontiguousarray(a, dtype=np.float64).tofile(f)
    r = subprocess.run([FORTRAN_CAS_SYM_BIN, str(fin), str(fout)], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    out = np.fromfile(fout, dtype=np.float64)
    nt = q["fsym0"].shape[1]
    assert out.size == 2 * (2 * nt + nw)
    for fi, fa, w in (np.split(half, [nt, 2 * nt]) for half in np.split(out, 2)):
        assert np.max(np.abs(fi - q["want_f"][0])) <= 1e-11 * np.max(np.abs(q["want_f"][0]))
        if do_active:
            assert np.max(np.abs(fa - q["want_f"][1])) <= 1e-11 * np.max(np.abs(q["want_f"][1]))
        else:
            assert np.array_equal(fa, q["fsym0"][1])
        assert np.max(np.abs(w - q["want_w"])) <= 1e-11 * np.max(np.abs(q["want_w"]))


@pytest.mark.skipif(not os.path.exists(FORTRAN_CHOINFO_BIN), reason="Fortran driver not built (no amdflang)")
def test_fortran_caller_reads_cholesky_files(gpu, tmp_path):
    import test_chofiles as tc
    from _chofiles import write_cholesky_files_sym
    from openmolcas_amd.fock import CholeskyFiles, JKContext

    sets, vec_sym, vec_set, vec_lt, diag = tc.sym_problem(seed=41, nvec=(8, 5, 4, 3))
    write_cholesky_files_sym(tmp_path, tc.NBAS_SYM, tc.ISOSHL_SYM, sets, vec_sym, vec_set, vec_lt, diag)
    nbas, nocc = tc.NBAS_SYM, [2, 1, 2, 1]
    rng = np.random.default_rng(42)
    porb = [rng.standard_normal((b, o)) for b, o in zip(nbas, nocc)]   # A2(nOcc(s), nBas(s)) memory
    ntri = sum(b * (b + 1) // 2 for b in nbas)
    dlt = rng.standard_normal(ntri)
    flt0 = rng.standard_normal(ntri)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(fin, "wb") as f:
        np.array([len(nbas)] + nbas + list(tc.ISOSHL_SYM) + nocc + [0.5], dtype=np.float64).tofile(f)
        for a in [np.concatenate([b.reshape(-1) for b in porb]), dlt, flt0]:
            np.ascontiguousarray(a, dtype=np.float64).tofile(f)
    r = subprocess.run([FORTRAN_CHOINFO_BIN, str(tmp_path), str(fin), str(fout)], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    got = np.fromfile(fout, dtype=np.float64)
    cf = CholeskyFiles(str(tmp_path / "CHORST"), str(tmp_path / "CHRED"), tc.ISOSHL_SYM)
    jk = JKContext(cf.nbas, cf.numcho, device=gpu)
    cf.load(jk, [str(tmp_path / f"CHVEC{s + 1}") for s in range(cf.nsym)])
    want = flt0.copy()[None]
    jk.cho_fscf_sym(nbas, 1, want, [[0] * 4], [nocc], [porb], dlt, 0.5)
    assert np.array_equal(got, want[0])
    assert np.max(np.abs(got - flt0)) > 0
    cf.close()
    jk.close()
