"""Pins the oracle's LK-screened exchange (mjk_oracle_lk.c) on the CPU.

* thresholds: the numbers of cho_lk_scf.F90:143-171 and
  cho_lk_screeningthreshold.F90:34-39 for chosen inputs, damping and floor;
* dmpk = 0 (every inequality trivially true) equals the unscreened
  CHO_LK_SCF / CHO_LK_CASSCF restatements;
* on shell-structured local inputs, the kept lists, half-transformed shell
  pairs and K blocks, and the screened K itself, equal an independent numpy
  evaluation of the reference's inequalities (tests/_lk_brute.py) across
  nScreen cadences, Estimate/Update switches, one and two densities and the
  SCF / CASSCF list cuts;
* two fragments far apart: every skipped contribution is exactly zero, so
  the screened result equals the unscreened one while work is skipped.
"""
import numpy as np
import pytest

import oracle
from _lk_brute import screened_exchange
from _lk_molecule import local_orbitals, molecule, store

THRCOM = 1e-6


def fold(a):
    n = a.shape[0]
    f = 2.0 * a
    f[np.diag_indices(n)] = np.diag(a)
    i, j = np.tril_indices(n)
    return f[i, j]


def rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def test_lk_thresholds_follow_the_reference_formulas():
    # LKThr = max(min(ThrCom, 1e-6) * dFmat, 1e-15); tau = LKThr/nocc*dmpk; thrv = sqrt(LKThr/(nocc*NumVT))*dmpk
    tau, thrv = oracle.lk_thresholds(1e-4, 0.5, 0.1, [4, 0], numvt=100)
    assert np.allclose(tau, [5e-7 / 4 * 0.1, 5e-7 / 1 * 0.1], rtol=1e-15)
    assert np.allclose(thrv, [np.sqrt(5e-7 / 400) * 0.1, np.sqrt(5e-7 / 100) * 0.1], rtol=1e-15)
    tau, _ = oracle.lk_thresholds(1e-8, 2.0, 1.0, [1], numvt=1)        # dFmat > 1: no damping
    assert tau[0] == pytest.approx(1e-8, rel=1e-15)
    tau, _ = oracle.lk_thresholds(1e-8, 1e-9, 1.0, [1], numvt=1)       # floor 1e-15
    assert tau[0] == pytest.approx(1e-15, rel=1e-15)
    tau, _ = oracle.lk_thresholds(1e-6, -1.0, 0.2, [2], numvt=5, estimate=True, maxred=4)
    assert tau[0] == pytest.approx(1e-6 / 2 * 0.2 / 4, rel=1e-15)     # Estimate: tau / MaxRed


def scf_case(mol, nocc, seed):
    Cs = [local_orbitals(mol["pos"], o, seed + d) for d, o in enumerate(nocc)]
    P = sum(C @ C.T for C in Cs)
    return Cs, fold(P)


@pytest.mark.parametrize("nden", [1, 2])
def test_lk_scf_dmpk0_equals_unscreened(nden):
    mol = molecule(natom=5, seed=3, nvec=40)
    st = store(mol)
    nocc = [6, 4][:nden]
    Cs, plt = scf_case(mol, nocc, 10)
    flt0 = np.random.default_rng(4).standard_normal((nden, npr := mol["n"] * (mol["n"] + 1) // 2))
    f_u, k_u = oracle.cho_lk_scf(st, flt0, nocc, Cs, plt, 0.5, nbassh=mol["nbassh"])
    f_s, k_s, stats = oracle.cho_lk_scf_screened(st, flt0, nocc, Cs, plt, 0.5, mol["nbassh"], mol["diag"],
                                                  THRCOM, nscreen=2, dmpk=0.0)
    assert rel(f_s, f_u) < 1e-12 and rel(k_s, k_u) < 1e-12
    assert npr > 0 and stats[2] > 0


def test_lk_casscf_dmpk0_equals_unscreened():
    mol = molecule(natom=5, seed=5, nvec=40)
    st = store(mol)
    n = mol["n"]
    rng = np.random.default_rng(6)
    cmo, _ = np.linalg.qr(rng.standard_normal((n, n)))
    Ci, Ca = cmo[:, :5], cmo[:, 5:9]
    occ = rng.uniform(0.2, 1.8, 4)
    dlt = np.stack([fold(2 * Ci @ Ci.T), fold((Ca * occ) @ Ca.T)])
    flt0 = rng.standard_normal((2, n * (n + 1) // 2))
    args = (np.sqrt(2) * Ci, -0.5, 5, 4, 4, Ca, Ca * np.sqrt(occ), True, cmo, 1, n - 1, 0.9)
    f_u, w_u = oracle.cho_lk_casscf(st, dlt, flt0, *args)
    f_s, w_s, _ = oracle.cho_lk_casscf_screened(st, dlt, flt0, *args, mol["nbassh"], mol["diag"], THRCOM,
                                                 dmpk=0.0)
    assert rel(f_s, f_u) < 1e-12 and rel(w_s, w_u) < 1e-13


@pytest.mark.parametrize("nden,nscreen,estimate,update,dmpk", [
    (1, 10, False, True, 0.1),
    (2, 1, False, True, 0.1),
    (1, 2, False, True, 1.0),
    (2, 2, True, True, 0.5),
    (1, 1, False, False, 0.3),
    (2, 0, False, True, 0.1),
])
def test_lk_scf_screened_matches_brute_force(nden, nscreen, estimate, update, dmpk):
    mol = molecule(natom=7, seed=11 + nden, nvec=80, nred=4)
    st = store(mol)
    nocc = [7, 5][:nden]
    Cs, plt = scf_case(mol, nocc, 20)
    n = mol["n"]
    z = np.zeros((nden, n * (n + 1) // 2))
    fj, _ = oracle.cho_lk_scf(st, z, [0] * nden, [np.zeros((n, 0))] * nden, plt, 0.5, nbassh=mol["nbassh"])
    f_s, _, stats = oracle.cho_lk_scf_screened(st, z, nocc, Cs, plt, 0.5, mol["nbassh"], mol["diag"], THRCOM,
                                                nscreen=nscreen, dmpk=dmpk, dfmat=1.0, estimate=estimate,
                                                update=update)
    tau, thrv = oracle.lk_thresholds(THRCOM, 1.0, dmpk, nocc, numvt=len(mol["vecs"]), estimate=estimate,
                                     maxred=len(mol["sets"]))
    K, bstats = screened_exchange(n, mol["sets"], mol["vred"], mol["vecs"], Cs, nocc, [-0.5] * nden, tau,
                                  thrv, mol["nbassh"], mol["diag"], nscreen, len(mol["sets"]), False,
                                  estimate, update)
    assert list(stats) == list(bstats)
    assert rel(f_s - fj, K) < 1e-12
    # screening is active: fewer K blocks than the unscreened count
    _, _, s0 = oracle.cho_lk_scf_screened(st, z, nocc, Cs, plt, 0.5, mol["nbassh"], mol["diag"], THRCOM,
                                          nscreen=nscreen, dmpk=0.0)
    assert stats[2] < s0[2] and stats[1] < s0[1]


@pytest.mark.parametrize("do_active,nscreen", [(True, 10), (True, 1), (False, 2)])
def test_lk_casscf_screened_matches_brute_force(do_active, nscreen):
    mol = molecule(natom=7, seed=31, nvec=75, nred=3)
    st = store(mol)
    n = mol["n"]
    Ci = local_orbitals(mol["pos"], 6, 32)
    Ca = local_orbitals(mol["pos"], 4, 33)
    occ = np.random.default_rng(34).uniform(0.2, 1.8, 4)
    msq, ash2 = np.sqrt(2) * Ci, Ca * np.sqrt(occ)
    dlt = np.stack([fold(msq @ msq.T), fold(ash2 @ ash2.T)])
    z = np.zeros((2, n * (n + 1) // 2))
    exfac, factxi, dmpk = 0.8, -0.5, 0.2
    args = (msq, factxi, 6, 0, 4, None, ash2, do_active, None, 0, 0, exfac)
    fj, _ = oracle.cho_lk_casscf(st, dlt, z, msq, factxi, 0, 0, 0, None, ash2, do_active, None, 0, 0, exfac)
    f_s, _, stats = oracle.cho_lk_casscf_screened(st, dlt, z, *args, mol["nbassh"], mol["diag"], THRCOM,
                                                  nscreen=nscreen, dmpk=dmpk)
    nden = 2 if do_active else 1
    nocc = [6, 4][:nden]
    tau, thrv = oracle.lk_thresholds(THRCOM, 1.0, dmpk, nocc, numvt=len(mol["vecs"]))
    K, bstats = screened_exchange(n, mol["sets"], mol["vred"], mol["vecs"], [msq, ash2][:nden], nocc,
                                  [factxi * exfac, -0.5 * exfac][:nden], tau, thrv, mol["nbassh"], mol["diag"],
                                  nscreen, len(mol["sets"]), True)
    assert list(stats) == list(bstats)
    assert rel((f_s - fj)[:nden], K) < 1e-12


def test_lk_disjoint_fragments_skip_only_zeros():
    """Two fragments 1000 units apart: every vector, orbital and density lives
    on one fragment, so every cross-fragment contribution is exactly zero and
    the screened build equals the unscreened one while skipping work."""
    a = molecule(natom=3, seed=41, nvec=30, nred=1)
    b = molecule(natom=3, seed=42, nvec=30, nred=1)
    na, nb = a["n"], b["n"]
    n = na + nb
    ia, ib = np.tril_indices(n)
    npr = len(ia)
    L = np.zeros((60, npr))
    lt = lambda i, j: i * (i + 1) // 2 + j  # noqa: E731
    for j in range(30):
        sqa = np.zeros((na, na))
        sqa[np.tril_indices(na)] = a["L"][j]
        sqb = np.zeros((nb, nb))
        sqb[np.tril_indices(nb)] = b["L"][j]
        full = np.zeros((n, n))
        full[:na, :na] = sqa
        L[2 * j] = full[ia, ib]
        full = np.zeros((n, n))
        full[na:, na:] = sqb
        L[2 * j + 1] = full[ia, ib]
    diag = np.sum(L * L, axis=0)
    st = oracle.CholStore.full(L)
    Ca = np.zeros((n, 3))
    Ca[:na] = local_orbitals(a["pos"], 3, 43)
    Cb = np.zeros((n, 2))
    Cb[na:] = local_orbitals(b["pos"], 2, 44)
    C = np.concatenate([Ca, Cb], axis=1)
    plt = fold(C @ C.T)
    nbassh = a["nbassh"] + b["nbassh"]
    z = np.zeros((1, npr))
    f_u, _ = oracle.cho_lk_scf(st, z, [5], [C], plt, 0.5, nbassh=nbassh)
    f_s, _, stats = oracle.cho_lk_scf_screened(st, z, [5], [C], plt, 0.5, nbassh, diag, 1e-14, dmpk=1e-6)
    _, _, s0 = oracle.cho_lk_scf_screened(st, z, [5], [C], plt, 0.5, nbassh, diag, 1e-14, dmpk=0.0)
    assert rel(f_s, f_u) < 1e-13
    assert stats[0] < s0[0] and stats[2] < s0[2]
    assert lt(1, 0) == 1
