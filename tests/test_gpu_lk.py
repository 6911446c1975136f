"""The reference's LK screening on the device (mjk_set_lk_screening with
mjk_set_shells / mjk_set_lk_diag) against the oracle's screened
CHO_LK_SCF / CHO_LK_CASSCF (oracle/mjk_oracle_lk.c, pinned in
tests/test_oracle_lk.py), FP64, relative 1e-11 on FLT, KLT and (pw|xy), and the
same counts of listed shells, half-transformed shell pairs and K blocks.

Inputs: shell-structured local vectors over three or four reduced sets of
shrinking pair lists (tests/_lk_molecule.py), localized orbitals; nScreen
cadences 1, 2, 10, 0; Estimate / Update; one and two densities; the SCF and
the CASSCF list cuts; dmpk = 0 against the unscreened drop-in."""
import numpy as np
import pytest

import oracle
from _lk_molecule import local_orbitals, molecule, store
from openmolcas_amd import MjkError
from openmolcas_amd.fock import JKContext

pytestmark = pytest.mark.gpu
RTOL = 1e-11
THRCOM = 1e-6


def fold(a):
    f = 2.0 * a
    f[np.diag_indices(a.shape[0])] = np.diag(a)
    i, j = np.tril_indices(a.shape[0])
    return f[i, j]


def rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def context(mol, gpu, estimate=False, update=True):
    n = mol["n"]
    jk = JKContext(n, len(mol["vecs"]), device=gpu)
    for r, s in enumerate(mol["sets"]):
        jk.set_reduced_set(r, s)
    vred = np.asarray(mol["vred"])
    for r in range(len(mol["sets"])):
        idx = np.nonzero(vred == r)[0]
        if len(idx):
            jk.upload_vectors(r, int(idx[0]), np.stack([mol["vecs"][j] for j in idx]))
    jk.set_shells(mol["nbassh"])
    jk.set_lk_diag(mol["diag"], maxred=len(mol["sets"]), estimate=estimate, update=update)
    jk.set_lk_screening(True, thrcom=THRCOM)
    return jk


@pytest.mark.parametrize("nden,nscreen,estimate,update,dmpk", [
    (1, 10, False, True, 0.1),
    (2, 1, False, True, 0.1),
    (1, 2, False, True, 1.0),
    (2, 2, True, True, 0.5),
    (1, 1, False, False, 0.3),
    (2, 0, False, True, 0.1),
])
def test_lk_scf_screened_matches_oracle(gpu, nden, nscreen, estimate, update, dmpk):
    mol = molecule(natom=8, seed=51 + nden, nvec=90, nred=4)
    n = mol["n"]
    nocc = [9, 6][:nden]
    Cs = [local_orbitals(mol["pos"], o, 60 + d) for d, o in enumerate(nocc)]
    plt = fold(sum(C @ C.T for C in Cs))
    rng = np.random.default_rng(7)
    flt0 = rng.standard_normal((nden, n * (n + 1) // 2))
    want_f, want_k, want_s = oracle.cho_lk_scf_screened(
        store(mol), flt0, nocc, Cs, plt, 0.5, mol["nbassh"], mol["diag"], THRCOM, nscreen=nscreen,
        dmpk=dmpk, dfmat=1.0, estimate=estimate, update=update)
    jk = context(mol, gpu, estimate, update)
    flt = [flt0[d].copy() for d in range(nden)]          # separately allocated, as the reference's
    klt = [np.zeros(n * (n + 1) // 2) for _ in range(nden)]
    jk.cho_lk_scf(nden, flt, klt, [0] * nden, nocc, Cs, plt, 0.5, nscreen=nscreen, dmpk=dmpk, dfmat=1.0,
                  nbassh=mol["nbassh"])
    st = jk.lk_stats()
    print("LK SCF", nden, nscreen, estimate, update, dmpk, "stats", st, "oracle", want_s)
    assert list(st[:3]) == list(want_s)
    assert rel(np.stack(flt), want_f) < RTOL
    assert rel(np.stack(klt), want_k) < RTOL
    jk.close()


@pytest.mark.parametrize("do_active,nscreen", [(True, 10), (True, 1), (False, 2)])
def test_lk_casscf_screened_matches_oracle(gpu, do_active, nscreen):
    mol = molecule(natom=8, seed=71, nvec=85, nred=3)
    n = mol["n"]
    rng = np.random.default_rng(72)
    cmo, _ = np.linalg.qr(rng.standard_normal((n, n)))
    Ci = local_orbitals(mol["pos"], 7, 73)
    Ca = local_orbitals(mol["pos"], 4, 74)
    occ = rng.uniform(0.2, 1.8, 4)
    msq, ash2 = np.sqrt(2) * Ci, Ca * np.sqrt(occ)
    dlt = np.stack([fold(msq @ msq.T), fold(ash2 @ ash2.T)])
    flt0 = rng.standard_normal((2, n * (n + 1) // 2))
    nfro, norb, exfac, dmpk = 1, n - 1, 0.8, 0.2
    want_f, want_w, want_s = oracle.cho_lk_casscf_screened(
        store(mol), dlt, flt0, msq, -0.5, 7, 4, 4, Ca, ash2, do_active, cmo, nfro, norb, exfac,
        mol["nbassh"], mol["diag"], THRCOM, nscreen=nscreen, dmpk=dmpk)
    jk = context(mol, gpu)
    flt = [flt0[0].copy(), flt0[1].copy()]
    w = np.zeros(norb * 4 * 10)
    jk.cho_lk_casscf([dlt[0].copy(), dlt[1].copy()], flt, msq, w, -0.5, 7, 4, 4, Ca, ash2, do_active, cmo,
                     nfro, norb, nscreen=nscreen, dmpk=dmpk, dfmat=1.0, exfac=exfac)
    st = jk.lk_stats()
    print("LK CASSCF", do_active, nscreen, "stats", st, "oracle", want_s)
    assert list(st[:3]) == list(want_s)
    nd = 2 if do_active else 1
    assert rel(np.stack(flt)[:nd], want_f[:nd]) < RTOL
    assert rel(w, want_w) < RTOL
    jk.close()


def test_lk_dmpk0_equals_unscreened_dropin(gpu):
    mol = molecule(natom=6, seed=81, nvec=60, nred=3)
    n = mol["n"]
    Cs = [local_orbitals(mol["pos"], 8, 82)]
    plt = fold(Cs[0] @ Cs[0].T)
    jk = context(mol, gpu)
    f_s = np.zeros((1, n * (n + 1) // 2))
    jk.cho_lk_scf(1, f_s, None, [0], [8], Cs, plt, 0.5, dmpk=0.0)
    jk.set_lk_screening(False)
    f_u = np.zeros((1, n * (n + 1) // 2))
    jk.cho_lk_scf(1, f_u, None, [0], [8], Cs, plt, 0.5, dmpk=0.0)
    assert rel(f_s, f_u) < 1e-12
    jk.close()


def test_lk_screening_needs_its_state(gpu):
    mol = molecule(natom=4, seed=91, nvec=30, nred=2)
    n = mol["n"]
    jk = JKContext(n, len(mol["vecs"]), device=gpu)
    jk.upload_packed(mol["L"])
    jk.set_lk_screening(True, thrcom=THRCOM)
    C = [local_orbitals(mol["pos"], 3, 92)]
    f = np.zeros((1, n * (n + 1) // 2))
    with pytest.raises(MjkError) as e:   # no diagonal
        jk.cho_lk_scf(1, f, None, [0], [3], C, fold(C[0] @ C[0].T), 0.5, nbassh=mol["nbassh"])
    assert e.value.rc == 3
    jk.set_lk_diag(mol["diag"])
    jk.cho_lk_scf(1, f, None, [0], [3], C, fold(C[0] @ C[0].T), 0.5, nbassh=mol["nbassh"])
    with pytest.raises(MjkError):          # screened builds are FP64
        jk.cho_lk_scf(1, f, None, [0], [3], C, fold(C[0] @ C[0].T), 0.5, precision="fp32")
    with pytest.raises(MjkError):          # shells must add up to nbas
        jk.set_shells([1] * (n - 1))
    jk.close()
