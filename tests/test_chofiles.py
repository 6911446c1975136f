"""The reader of the reference's Cholesky files (mjk_choinfo_read, host code:
no device needed) against files written in the reference's layout by
tests/_chofiles.py: dimensions, ThrCom, InfVec, and the LT map of every
reduced set rebuilt from CHRED's shell-pair-local indices and ISOSHL, for
shells whose AOs are consecutive and for shells whose AOs are interleaved
(the way p and d components of one centre interleave); with four irreps,
every (irrep, reduced set) map against the writer and against a forward
rebuild that follows CHO_RSTOF / CHO_RS2F on the raw CHRED words; error codes
for a missing file, nSym outside 1..8, an element filed under the wrong
irrep and an ISOSHL out of range."""
import os

import numpy as np
import pytest

from _chofiles import itri, write_cholesky_files, write_cholesky_files_sym
from openmolcas_amd import MjkError, _lib
from openmolcas_amd.fock import CholeskyFiles

pytestmark = pytest.mark.skipif(not os.path.exists(_lib.LIB_PATH), reason="library not built")


def problem(n, isoshl, nvec, seed, nred=3):
    rng = np.random.default_rng(seed)
    npr = n * (n + 1) // 2
    s0 = np.sort(rng.choice(npr, size=int(0.8 * npr), replace=False))
    sets = [s0]
    for r in range(1, nred):
        sets.append(np.sort(rng.choice(sets[-1], size=int(0.6 * len(sets[-1])), replace=False)))
    vec_set = np.sort(rng.integers(0, nred, nvec))
    vec_lt = []
    for J in range(nvec):
        v = np.zeros(npr)
        v[sets[vec_set[J]]] = rng.standard_normal(len(sets[vec_set[J]]))
        vec_lt.append(v)
    diag = np.zeros(npr)
    diag[s0] = rng.uniform(0.1, 2.0, len(s0))
    return sets, vec_set, vec_lt, diag


@pytest.mark.parametrize("layout,adrvec", [("contiguous", 1), ("interleaved", 2), ("interleaved", 1)])
def test_reader_rebuilds_reduced_sets(tmp_path, layout, adrvec):
    if layout == "contiguous":
        sizes = [1, 3, 1, 5, 3, 1]
        isoshl = np.repeat(np.arange(1, len(sizes) + 1), sizes)
    else:  # s, p(x,y,z) and d components of two shells interleaved
        isoshl = np.array([1, 2, 3, 2, 3, 2, 3, 4, 5, 4, 5, 4, 5, 4, 5, 4, 5, 6])
    n = len(isoshl)
    sets, vec_set, vec_lt, diag = problem(n, isoshl, 17, seed=5)
    maps, vecs, iadr = write_cholesky_files(tmp_path, n, isoshl, sets, vec_set, vec_lt, diag,
                                            adrvec=adrvec, thrcom=3e-7)
    cf = CholeskyFiles(str(tmp_path / "CHORST"), str(tmp_path / "CHRED"), isoshl)
    assert (cf.nbas, cf.numcho, cf.maxred, cf.adrvec) == (n, 17, 3, adrvec)
    assert cf.nshell == int(isoshl.max()) and cf.thrcom == 3e-7
    for r in range(3):
        got = cf.reduced_set(r)
        assert np.array_equal(got, maps[r])           # same elements in file order
        assert np.array_equal(np.sort(got), np.sort(sets[r]))
    assert np.array_equal(cf.infvec(2), vec_set + 1)
    assert np.array_equal(cf.infvec(3), iadr)
    cf.close()


def test_reader_errors(tmp_path):
    isoshl = np.array([1, 2, 2, 2])
    sets, vec_set, vec_lt, diag = problem(4, isoshl, 3, seed=6, nred=1)
    write_cholesky_files(tmp_path, 4, isoshl, sets, vec_set, vec_lt, diag)
    with pytest.raises(MjkError) as e:
        CholeskyFiles(str(tmp_path / "nope"), str(tmp_path / "CHRED"), isoshl)
    assert e.value.rc == 4
    with pytest.raises(MjkError) as e:
        CholeskyFiles(str(tmp_path / "CHORST"), str(tmp_path / "CHRED"), np.array([1, 2, 3, 2]))
    assert e.value.rc == 1                            # shell 3 > nShell
    raw = np.fromfile(tmp_path / "CHORST", dtype="<i8")
    raw[0] = 9                                        # nine irreps: no point group has them
    raw.tofile(tmp_path / "CHORST2")
    with pytest.raises(MjkError) as e:
        CholeskyFiles(str(tmp_path / "CHORST2"), str(tmp_path / "CHRED"), isoshl)
    assert e.value.rc == 1


# four irreps (D2 / C2v); shells spread over irreps the way symmetry-adapted
# combinations of one shell fall into several irreps
NBAS_SYM = [5, 3, 4, 2]
ISOSHL_SYM = np.array([1, 2, 3, 4, 4, 2, 3, 4, 2, 3, 5, 5, 4, 5])


def sym_problem(seed, nred=3, nvec=(5, 3, 2, 2)):
    rng = np.random.default_rng(seed)
    n = sum(NBAS_SYM)
    so_sym = np.repeat(np.arange(len(NBAS_SYM)), NBAS_SYM)
    i, j = np.tril_indices(n)
    pair_sym = so_sym[i] ^ so_sym[j]
    npr = n * (n + 1) // 2
    s0 = np.sort(rng.choice(npr, size=int(0.8 * npr), replace=False))
    sets = [s0]
    for r in range(1, nred):
        sets.append(np.sort(rng.choice(sets[-1], size=int(0.6 * len(sets[-1])), replace=False)))
    vec_sym, vec_set, vec_lt = [], [], []
    for sym, nv in enumerate(nvec):
        for r in np.sort(rng.integers(0, nred, nv)):
            v = np.zeros(npr)
            on = sets[r][pair_sym[sets[r]] == sym]
            v[on] = rng.standard_normal(len(on))
            vec_sym.append(sym)
            vec_set.append(int(r))
            vec_lt.append(v)
    diag = np.zeros(npr)
    diag[s0] = rng.uniform(0.1, 2.0, len(s0))
    return sets, vec_sym, vec_set, vec_lt, diag


def rstof(d, nbas_sym, isoshl, maxred, ired):
    """IRS2F of reduced set ired (1-based) from the raw files, built forward the
    way CHO_RSTOF does (src/cholesky_util/cho_rstof.F90:34-89): every SO pair
    (a, b) of irreps symA >= symB, its shell-pair-local index LAB, and its
    position found by CHO_RS2F's search (cho_rs2f.F90: IndRed = LAB in set 1,
    IndRed(IndRed(K)) = LAB in later sets) inside the block of irrep
    symA x symB and the pair's shell pair."""
    nsym = len(nbas_sym)
    rst = np.fromfile(f"{d}/CHORST", dtype="<i8")
    red = np.fromfile(f"{d}/CHRED", dtype="<i8")
    nnshl = int(rst[2])
    infred = rst[4 + nsym + 2 + 8 + 1:4 + nsym + 2 + 8 + 1 + maxred]

    def block(r):
        a0 = int(infred[r - 1])
        nn = red[a0:a0 + nsym * nnshl].reshape(nnshl, nsym).T  # nnBstRSh(iSym, iShlAB)
        ind = red[a0 + nsym * nnshl:a0 + nsym * nnshl + int(nn.sum())]
        return nn, ind

    nn1, ind1 = block(1)
    n1 = int(nn1.sum())
    isp2f = red[nsym * nnshl + 2 * n1:nsym * nnshl + 2 * n1 + nnshl]
    f2sp = {int(f): k + 1 for k, f in enumerate(isp2f)}
    nn, ind = block(ired)
    iibstr = np.concatenate([[0], np.cumsum(nn.sum(axis=1))[:-1]])      # cho_setredind.F90
    iibstrsh = np.concatenate([np.zeros((nsym, 1), dtype=np.int64), np.cumsum(nn, axis=1)[:, :-1]], axis=1)
    nbstsh = np.bincount(isoshl)[1:]
    ishlso = np.zeros(len(isoshl), dtype=np.int64)
    c = {}
    for i, s in enumerate(isoshl):
        c[s] = c.get(s, 0) + 1
        ishlso[i] = c[s]
    ibas = np.concatenate([[0], np.cumsum(nbas_sym)[:-1]])
    out = {}
    for sa in range(nsym):
        for sb in range(sa + 1):
            sab = sa ^ sb
            for ka in range(nbas_sym[sa]):
                for kb in range(nbas_sym[sb] if sb < sa else ka + 1):
                    ia, ib = ibas[sa] + ka, ibas[sb] + kb
                    la, lb, ha, hb = ishlso[ia], ishlso[ib], isoshl[ia], isoshl[ib]
                    if ha < hb:
                        lab = nbstsh[hb - 1] * (la - 1) + lb
                    elif ha == hb:
                        lab = itri(la, lb)
                    else:
                        lab = nbstsh[ha - 1] * (lb - 1) + la
                    sp = f2sp.get(itri(ha, hb), 0)
                    if not sp:
                        continue
                    k0 = iibstr[sab] + iibstrsh[sab, sp - 1]
                    for k in range(k0, k0 + nn[sab, sp - 1]):
                        hit = ind[k] if ired == 1 else ind1[ind[k] - 1]
                        if hit == lab:
                            out[k] = (ia, ib)
                            break
    return out


@pytest.mark.parametrize("adrvec", [1, 2])
def test_reader_with_symmetry(tmp_path, adrvec):
    sets, vec_sym, vec_set, vec_lt, diag = sym_problem(seed=8)
    maps, vecs, iadrs = write_cholesky_files_sym(tmp_path, NBAS_SYM, ISOSHL_SYM, sets, vec_sym, vec_set, vec_lt,
                                                 diag, adrvec=adrvec, thrcom=2e-6)
    cf = CholeskyFiles(str(tmp_path / "CHORST"), str(tmp_path / "CHRED"), ISOSHL_SYM)
    assert (cf.nsym, cf.nbas, cf.maxred, cf.adrvec) == (4, 14, 3, adrvec)
    assert cf.nbas_sym == NBAS_SYM and cf.numcho_sym == [5, 3, 2, 2] and cf.numcho == 12
    for sym in range(4):
        for r in range(3):
            assert np.array_equal(cf.reduced_set(r, sym + 1), maps[(sym, r)]), (sym, r)
    # the forward rebuild of CHO_RSTOF gives the same element -> SO pair map
    for r in range(3):
        irs2f = rstof(tmp_path, NBAS_SYM, ISOSHL_SYM, 3, r + 1)
        got = np.concatenate([cf.reduced_set(r, sym + 1) for sym in range(4)])
        assert len(irs2f) == len(got)
        for k, (ia, ib) in irs2f.items():
            a, b = max(ia, ib), min(ia, ib)
            assert got[k] == a * (a + 1) // 2 + b
    assert np.array_equal(cf.infvec(2), np.array([3 * s + r + 1 for s, r in zip(vec_sym, vec_set)]))
    assert np.array_equal(cf.infvec(3), np.concatenate(iadrs))
    cf.close()


def test_reader_rejects_misfiled_irrep(tmp_path):
    sets, vec_sym, vec_set, vec_lt, diag = sym_problem(seed=9, nred=1)
    write_cholesky_files_sym(tmp_path, NBAS_SYM, ISOSHL_SYM, sets, vec_sym, vec_set, vec_lt, diag)
    red = np.fromfile(tmp_path / "CHRED", dtype="<i8")
    rst = np.fromfile(tmp_path / "CHORST", dtype="<i8")
    nnshl = int(rst[2])
    nn = red[:4 * nnshl].reshape(nnshl, 4).copy()
    assert not np.array_equal(nn[:, 0], nn[:, 1])
    nn[:, [0, 1]] = nn[:, [1, 0]]                     # irreps 1 and 2 swap their element counts
    red[:4 * nnshl] = nn.reshape(-1)
    red.tofile(tmp_path / "CHRED")
    with pytest.raises(MjkError) as e:
        CholeskyFiles(str(tmp_path / "CHORST"), str(tmp_path / "CHRED"), ISOSHL_SYM)
    assert e.value.rc == 1 and "irrep" in str(e.value)
