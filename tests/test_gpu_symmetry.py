"""CHO_FSCF and CHO_LK_SCF with point-group symmetry (mjk_cho_fscf_sym,
mjk_cho_lk_scf_sym) against the oracle's C1 CHO_FSCF on the same problem
written in the full basis (unscreened LK equals CHO_FSCF with ExFac = FactXI).
Two irreps (nbas 7 + 4), vectors of JSYM = 1 and 2 in their own reduced sets
(tests/_sym.py).  Tolerance: relative 1e-11 of max|FLT| (FP64 throughout)."""
import numpy as np
import pytest

from _sym import NBAS, make_casscf_problem, make_problem
from openmolcas_amd import MjkError
from openmolcas_amd.fock import JKContext

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("lk", [False, True])
@pytest.mark.parametrize("nden,exfac", [(1, 0.5), (2, 1.0)])
def test_cho_fscf_symmetry_matches_c1_oracle(gpu, nden, exfac, lk):
    p = make_problem(nden, exfac, seed=7 + nden)
    # the problem is block diagonal: the C1 result has no off-irrep entries
    assert np.max(np.abs(p["want_full"][:, p["offblock"]])) <= 1e-12 * np.max(np.abs(p["want_full"]))
    m1 = p["V1"].shape[0]
    jk = JKContext(p["n"], m1 + p["V2"].shape[0], device=gpu)
    jk.set_reduced_set(0, np.array(p["rs"][0], dtype=np.int64))
    jk.set_reduced_set(1, np.array(p["rs"][1], dtype=np.int64))
    jk.upload_vectors(0, 0, p["V1"])
    jk.upload_vectors(1, m1, p["V2"])
    got = p["fsym"].copy()
    if lk:
        jk.cho_lk_scf_sym(NBAS, nden, got, p["nforb"], p["niorb"], p["porb"], p["dsym"], exfac,
                          dmpk=0.1, dfmat=1.0)
    else:
        jk.cho_fscf_sym(NBAS, nden, got, p["nforb"], p["niorb"], p["porb"], p["dsym"], exfac)
    jk.close()
    want = p["want"]
    assert np.max(np.abs(got - want)) <= 1e-11 * np.max(np.abs(want))



@pytest.mark.parametrize("do_active", [True, False])
def test_cho_fmcscf_symmetry_matches_c1_oracle(gpu, do_active):
    """CHO_FMCSCF, nSym = 2: FI/FA irrep blocks against the oracle's C1
    CHO_FMCSCF on the embedded problem; W_PWXY against (pw|xy) formed from the
    dense vectors and MOs, laid out in blocks (sp, sw, sx <= sy) with
    sp x sw = sx x sy (cho_eval_waxy.F90:86-141).  Tolerance: relative 1e-11."""
    q = make_casscf_problem(do_active, seed=31)
    p = q["vectors"]
    jk = JKContext(p["n"], p["V1"].shape[0] + p["V2"].shape[0], device=gpu)
    jk.set_reduced_set(0, np.array(p["rs"][0], dtype=np.int64))
    jk.set_reduced_set(1, np.array(p["rs"][1], dtype=np.int64))
    jk.upload_vectors(0, 0, p["V1"])
    jk.upload_vectors(1, p["V1"].shape[0], p["V2"])
    fl = q["fsym0"].copy()
    w = np.full(q["want_w"].size, np.nan)
    jk.cho_fmcscf_sym(NBAS, fl, q["nforb"], q["niorb"], q["naorb"], q["factxi"], q["dsym"], do_active,
                      q["porb1"], q["nchm"], q["porb2"], q["porb3"], w, q["cmo"], q["nfro"], q["norb"])
    jk.close()
    nd = 2 if do_active else 1
    want = q["want_f"][:nd]
    assert np.max(np.abs(fl[:nd] - want)) <= 1e-11 * np.max(np.abs(want))
    if not do_active:
        assert np.array_equal(fl[1], q["fsym0"][1])  # FA untouched
    assert np.max(np.abs(w - q["want_w"])) <= 1e-11 * np.max(np.abs(q["want_w"]))


@pytest.mark.parametrize("do_active", [True, False])
def test_cho_lk_casscf_symmetry_matches_c1_oracle(gpu, do_active):
    """CHO_LK_CASSCF, nSym = 2, unscreened with ExFac = 1: the same FI, FA and
    W_PWXY as CHO_FMCSCF on that problem (MSQ columns = POrb(1), CVa(2)
    columns = POrb(2)); LK screening with nSym > 1 is refused."""
    q = make_casscf_problem(do_active, seed=37)
    p = q["vectors"]
    jk = JKContext(p["n"], p["V1"].shape[0] + p["V2"].shape[0], device=gpu)
    jk.set_reduced_set(0, np.array(p["rs"][0], dtype=np.int64))
    jk.set_reduced_set(1, np.array(p["rs"][1], dtype=np.int64))
    jk.upload_vectors(0, 0, p["V1"])
    jk.upload_vectors(1, p["V1"].shape[0], p["V2"])
    nchi = [q["nforb"][s] + q["niorb"][s] for s in range(2)]
    fl = q["fsym0"].copy()
    w = np.full(q["want_w"].size, np.nan)
    jk.cho_lk_casscf_sym(NBAS, q["dsym"], fl, q["porb1"], w, q["factxi"], nchi, q["naorb"], q["nchm"],
                         q["porb3"], q["porb2"], do_active, q["cmo"], q["nfro"], q["norb"])
    nd = 2 if do_active else 1
    want = q["want_f"][:nd]
    assert np.max(np.abs(fl[:nd] - want)) <= 1e-11 * np.max(np.abs(want))
    if not do_active:
        assert np.array_equal(fl[1], q["fsym0"][1])
    assert np.max(np.abs(w - q["want_w"])) <= 1e-11 * np.max(np.abs(q["want_w"]))
    jk.set_lk_screening(True)
    with pytest.raises(MjkError, match="nSym > 1"):
        jk.cho_lk_casscf_sym(NBAS, q["dsym"], fl, q["porb1"], w, q["factxi"], nchi, q["naorb"], q["nchm"],
                             q["porb3"], q["porb2"], do_active, q["cmo"], q["nfro"], q["norb"])
    jk.close()


def _fold_klt(klt_sym, nbassh):
    """Adds KLT(d)%SB(s) to a zero FLT the way CHO_LK_SCF folds it
    (cho_lk_scf.F90:963-1028): shell pairs iaSh > ibSh rectangular with ia
    fastest, diagonal shell pairs lower-triangular, block offsets accumulated
    over the shell pairs (cho_lk_scf.F90:274-292)."""
    out = np.zeros_like(klt_sym)
    toff = 0
    for s, nb in enumerate(NBAS):
        sh = list(nbassh[s])
        koff = np.concatenate([[0], np.cumsum(sh)])
        shp_off, acc = {}, 0
        for ia in range(len(sh)):
            for ib in range(ia + 1):
                shp_off[(ia, ib)] = acc
                acc += sh[ia] * sh[ib] - (sh[ia] * (sh[ia] - 1) // 2 if ia == ib else 0)
        for ia_sh in range(len(sh)):
            for ib_sh in range(ia_sh):
                o = shp_off[(ia_sh, ib_sh)]
                for ib in range(sh[ib_sh]):
                    for ia in range(sh[ia_sh]):
                        ag, bg = koff[ia_sh] + ia, koff[ib_sh] + ib
                        out[toff + ag * (ag + 1) // 2 + bg] += klt_sym[toff + o + sh[ia_sh] * ib + ia]
            o = shp_off[(ia_sh, ia_sh)]
            for ib in range(sh[ia_sh]):
                for ia in range(ib, sh[ia_sh]):
                    ag, bg = koff[ia_sh] + ia, koff[ia_sh] + ib
                    out[toff + ag * (ag + 1) // 2 + bg] += klt_sym[toff + o + ia * (ia + 1) // 2 + ib]
        toff += nb * (nb + 1) // 2
    return out


@pytest.mark.parametrize("nden", [1, 2])
def test_cho_lk_scf_symmetry_klt(gpu, nden):
    """KLT of CHO_LK_SCF with two irreps in each irrep's shell-pair storage:
    folded back the reference's way it equals -FactXI K of the oracle's C1
    build restricted to the irrep blocks (and FLT is unchanged by asking for
    it)."""
    import oracle

    p = make_problem(nden, 0.7, seed=17 + nden)
    n, factxi = p["n"], 0.7
    nbassh = [[3, 4], [1, 3]]  # nBasSh(iSym, iSh)
    st = oracle.CholStore(n, [np.array(p["rs"][0]), np.array(p["rs"][1])],
                          [(0, v) for v in p["V1"]] + [(1, v) for v in p["V2"]])
    cols = []  # the orbitals in the full basis, as make_problem builds them
    off = [0, NBAS[0]]
    for d in range(nden):
        c = []
        for s in range(2):
            for k in range(p["porb"][d][s].shape[1]):
                v = np.zeros(n)
                v[off[s]:off[s] + NBAS[s]] = p["porb"][d][s][:, k]
                c.append(v)
        cols.append(np.stack(c, axis=1))
    ii, jj = np.tril_indices(n)
    dfull = np.zeros(n * (n + 1) // 2)
    gidx = np.array([(off[s] + a) * (off[s] + a + 1) // 2 + off[s] + b
                     for s in range(2) for a in range(NBAS[s]) for b in range(a + 1)])
    dfull[gidx] = p["dsym"]
    _, kfull = oracle.cho_lk_scf(st, np.zeros((nden, n * (n + 1) // 2)), [c.shape[1] for c in cols], cols,
                                 dfull, factxi)
    want = kfull[:, gidx]
    jk = JKContext(n, p["V1"].shape[0] + p["V2"].shape[0], device=gpu)
    jk.set_reduced_set(0, np.array(p["rs"][0], dtype=np.int64))
    jk.set_reduced_set(1, np.array(p["rs"][1], dtype=np.int64))
    jk.upload_vectors(0, 0, p["V1"])
    jk.upload_vectors(1, p["V1"].shape[0], p["V2"])
    f1, f2 = p["fsym"].copy(), p["fsym"].copy()
    klt = np.zeros_like(p["fsym"])
    jk.cho_lk_scf_sym(NBAS, nden, f1, p["nforb"], p["niorb"], p["porb"], p["dsym"], factxi)
    jk.cho_lk_scf_sym(NBAS, nden, f2, p["nforb"], p["niorb"], p["porb"], p["dsym"], factxi,
                      klt=klt, nbassh=nbassh)
    jk.close()
    assert np.array_equal(f1, f2)
    for d in range(nden):
        got = _fold_klt(klt[d], nbassh)
        assert np.max(np.abs(got - want[d])) <= 1e-11 * np.max(np.abs(want[d]))
