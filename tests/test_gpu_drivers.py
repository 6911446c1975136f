"""The callers' dispatch (ChoSCF_Drv, CHO_CAS_DRV) on the device drop-ins.

Against dense numpy brute force written from the mathematics:
  RHF  F = J[D] - 1/2 ExFac K[D]                  (D doubly occupied)
  UHF  F_s = J[Da + Db] - ExFac K[D_s]            (s = alpha, beta)
  CAS  FI = J[DI] - 1/2 K[DI], FA = J[DA] - 1/2 K[DA], W = sum_J (pw|J)(xy|J)
with K[D] = sum_J L^J D L^J.  DECO and plain-MO routes, ALGO 3 and 4, must all
give these to relative 1e-10 (DECO truncates at 1e-8 of the max diagonal for
SCF, which for exact-rank densities removes nothing).
"""
import numpy as np
import pytest

from _dense import coulomb, lfull
from openmolcas_amd.drivers import cho_cas_drv, choscf_fock
from openmolcas_amd.fock import JKContext
from openmolcas_amd.synthetic import chol_vectors_np, pack

pytestmark = pytest.mark.gpu

RTOL = 1e-10


def rel(a, b):
    return np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(1.0, np.max(np.abs(b)))


def kmat(Lf, D):
    return np.einsum("jab,bc,jcd->ad", Lf, D, Lf)


@pytest.fixture(scope="module")
def system(gpu):
    n, M = 60, 90
    L = chol_vectors_np(n, M, seed=80)
    jk = JKContext(n, M, device=gpu)
    jk.upload_packed(L)
    cmo, _ = np.linalg.qr(np.random.default_rng(81).standard_normal((n, n)))
    yield jk, lfull(L, n), cmo
    jk.close()


@pytest.mark.parametrize("algo,deco,exfac", [(3, True, 1.0), (3, False, 1.0), (4, True, 0.25),
                                             (4, False, 1.0), (3, True, 0.0)])
def test_choscf_rhf(system, algo, deco, exfac):
    jk, Lf, cmo = system
    nocc = 9
    D = 2.0 * cmo[:, :nocc] @ cmo[:, :nocc].T
    F = choscf_fock(jk, D, exfac=exfac, algo=algo, deco=deco, cmo=cmo, nocc=nocc,
                    nbassh=[7, 13, 20, 20] if algo == 4 else None)
    J, _ = coulomb(Lf, D)
    want = pack(J - 0.5 * exfac * kmat(Lf, D))
    assert F.shape == (1, want.size)
    assert rel(F[0], want) < RTOL


@pytest.mark.parametrize("algo,deco", [(3, True), (3, False), (4, True), (4, False)])
def test_choscf_uhf(system, algo, deco):
    jk, Lf, cmo = system
    na, nb = 9, 6
    Da = cmo[:, :na] @ cmo[:, :na].T
    Db = cmo[:, :nb] @ cmo[:, :nb].T
    F = choscf_fock(jk, Da, exfac=1.0, algo=algo, deco=deco, cmo=cmo, nocc=na, dsq_beta=Db,
                    cmo_beta=cmo, nocc_beta=nb)
    J, _ = coulomb(Lf, Da + Db)
    assert rel(F[0], pack(J - kmat(Lf, Da))) < RTOL
    assert rel(F[1], pack(J - kmat(Lf, Db))) < RTOL


@pytest.mark.parametrize("deco,lk", [(True, False), (False, False), (True, True), (False, True)])
def test_cho_cas_drv(system, deco, lk):
    jk, Lf, cmo = system
    n = cmo.shape[0]
    nfro, nish, nash = 2, 5, 4
    nocc = nfro + nish
    Ci = cmo[:, :nocc]
    Ca = cmo[:, nocc:nocc + nash]
    rng = np.random.default_rng(82)
    q, _ = np.linalg.qr(rng.standard_normal((nash, nash)))
    rdm = (q * rng.uniform(0.1, 1.9, nash)) @ q.T  # active 1-RDM, eigenvalues in (0, 2)
    DI = 2.0 * Ci @ Ci.T
    DA = Ca @ rdm @ Ca.T
    FI, FA, W = cho_cas_drv(jk, cmo, DI, DA, nfro, nish, nash, deco=deco, lk=lk)
    JI, _ = coulomb(Lf, DI)
    JA, _ = coulomb(Lf, DA)
    assert rel(FI, pack(JI - 0.5 * kmat(Lf, DI))) < RTOL
    assert rel(FA, pack(JA - 0.5 * kmat(Lf, DA))) < RTOL
    norb = n - nfro
    pwJ = np.einsum("jab,ap,bw->jpw", Lf, cmo[:, nfro:], Ca)
    xyJ = np.einsum("jab,ax,by->jxy", Lf, Ca, Ca)
    nxy = nash * (nash + 1) // 2
    want = np.zeros((nxy, nash, norb))
    for y in range(nash):
        for x in range(y + 1):
            want[y * (y + 1) // 2 + x] = np.einsum("jpw,j->wp", pwJ, xyJ[:, x, y])
    assert rel(W, want.reshape(-1)) < RTOL
