"""Coulomb pass fused into the FP64 half-transform (mjk_coulomb_exchange).

J and V against the oracle's Coulomb build (FP64, relative 1e-11) and against
the two-pass mjk_coulomb; K bit-identical to mjk_exchange (the fused kernel
adds independent work to the same MFMA accumulation).  Densities unrelated to
the orbitals, so V can only come from D.  Cases: tile boundaries (n = 37, 129,
200, 730: a packed triangle of 2 MB, larger than an L2 share), one and two
orbital sets, a wide bank (12 column tiles, the 2 x 2 wave layout), 128-wide
AO tiles, several vector chunks (small X budget), a set without orbitals, a
reduced set without pairs, the CHO_FSCF drop-in that routes through the fused
path, and the argument errors."""
import numpy as np
import pytest
import torch

import oracle
from openmolcas_amd.fock import JKContext
from openmolcas_amd.synthetic import chol_vectors_np, fold, npair, orbitals_np

pytestmark = pytest.mark.gpu

RTOL = 1e-11


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.max(np.abs(a - b)) / max(1.0, np.max(np.abs(b)))


def random_density(n, seed):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, n))
    return fold((A + A.T) / (2.0 * n))


def oracle_jk(L, dlt, Cs, n):
    st = oracle.CholStore.full(L)
    z = np.zeros((1, npair(n)))
    J = oracle.cho_fscf(st, z, [0], [np.zeros((n, 0))], dlt, 0.0)[0]
    K = [oracle.cho_fscf(st, z, [c.shape[1]], [c], np.zeros(npair(n)), -1.0)[0] for c in Cs]
    return J, K


def oracle_v(L, dlt):
    return L @ dlt


def check(jk, L, dlt, Cs, n):
    Ct = [torch.from_numpy(c) for c in Cs]
    Dt = torch.from_numpy(dlt)
    J, K, V = jk.coulomb_exchange(Dt, Ct, want_v=True)
    J, K, V = J.cpu().numpy(), K.cpu().numpy(), V.cpu().numpy()[:, 0]
    Jw, Kw = oracle_jk(L, dlt, Cs, n)
    assert rel(J[0], Jw) < RTOL
    for d in range(len(Cs)):
        assert rel(K[d], Kw[d]) < RTOL
    assert rel(V, oracle_v(L, dlt)) < RTOL
    # against the two-pass Coulomb and the plain FP64 exchange
    J2, V2 = jk.coulomb(Dt.unsqueeze(0), want_v=True)
    assert rel(J[0], J2.cpu().numpy()[0]) < 1e-13
    assert rel(V, V2.cpu().numpy()[:, 0]) < 1e-13
    K2 = jk.exchange(Ct).cpu().numpy()
    assert np.array_equal(K, K2)


@pytest.mark.parametrize("n,M,occs", [(37, 21, (5,)), (129, 17, (1,)), (200, 33, (40, 9)),
                                      (70, 9, (190,)), (256, 40, (64,)), (730, 7, (20,)),
                                      (150, 13, (70,)), (170, 11, (100,))])
def test_fused_matches_oracle(gpu, n, M, occs):
    L = chol_vectors_np(n, M, seed=n + M)
    Cs = [orbitals_np(n, o, seed=o + 3) for o in occs]
    dlt = random_density(n, seed=n)
    jk = JKContext(n, M, device=gpu)
    jk.upload_packed(L)
    check(jk, L, dlt, Cs, n)
    jk.close()


def test_fused_wide_ao_tiles_and_chunks(gpu):
    n, M = 160, 50
    L = chol_vectors_np(n, M, seed=5)
    Cs = [orbitals_np(n, 48, seed=6)]
    dlt = random_density(n, seed=7)
    jk = JKContext(n, M, device=gpu)
    jk.upload_packed(L)
    jk.set_tuning(0, 0, 128)          # 128-wide AO tiles of the half-transform
    check(jk, L, dlt, Cs, n)
    jk.set_tuning(0, 0, 0)
    jk.set_x_budget(8 * 48 * 256 * 7)  # 7 vectors of X (48 x 256): 8 half-transform launches
    check(jk, L, dlt, Cs, n)
    jk.close()


def test_fused_set_without_orbitals(gpu):
    n, M = 45, 12
    L = chol_vectors_np(n, M, seed=8)
    dlt = random_density(n, seed=9)
    jk = JKContext(n, M, device=gpu)
    jk.upload_packed(L)
    C = orbitals_np(n, 6, seed=10)
    J, K = jk.coulomb_exchange(torch.from_numpy(dlt), [torch.zeros((n, 0), dtype=torch.float64),
                                                       torch.from_numpy(C)])
    Jw, Kw = oracle_jk(L, dlt, [C], n)
    assert rel(J.cpu().numpy()[0], Jw) < RTOL
    assert torch.count_nonzero(K[0]).item() == 0
    assert rel(K.cpu().numpy()[1], Kw[0]) < RTOL
    # no orbitals at all: the two-pass Coulomb carries J
    J0, K0 = jk.coulomb_exchange(torch.from_numpy(dlt), [torch.zeros((n, 0), dtype=torch.float64)])
    assert rel(J0.cpu().numpy()[0], Jw) < RTOL
    assert torch.count_nonzero(K0).item() == 0
    jk.close()


def test_fused_reduced_sets(gpu):
    n, M1 = 20, 6
    npr = npair(n)
    L0 = chol_vectors_np(n, M1, seed=4)
    jk = JKContext(n, M1 + 2, device=gpu)
    jk.set_reduced_set(0, np.arange(npr))
    jk.set_reduced_set(1, np.zeros(0, dtype=np.int64))
    jk.upload_vectors(0, 0, L0)
    jk.upload_vectors(1, M1, np.zeros((2, 0)))
    L = np.vstack([L0, np.zeros((2, npr))])
    check(jk, L, random_density(n, seed=11), [orbitals_np(n, 4, seed=5)], n)
    jk.close()


def test_cho_fscf_dropin_fused(gpu):
    """The FP64 drop-in now takes the fused path: FLT against the oracle for
    RHF and UHF-style two densities."""
    n, M = 150, 60
    L = chol_vectors_np(n, M, seed=12)
    Ca, Cb = orbitals_np(n, 11, seed=13), orbitals_np(n, 8, seed=14)
    dlt = fold(Ca @ Ca.T + Cb @ Cb.T)
    flt0 = np.random.default_rng(15).standard_normal((2, npair(n)))
    st = oracle.CholStore.full(L)
    want = oracle.cho_fscf(st, flt0, [11, 8], [Ca, Cb], dlt, 1.0)
    jk = JKContext(n, M, device=gpu)
    jk.upload_packed(L)
    got = flt0.copy()
    jk.cho_fscf(2, got, [0, 0], [11, 8], [Ca, Cb], dlt, 1.0)
    assert rel(got, want) < RTOL
    jk.close()


def test_fused_argument_errors(gpu):
    """Error paths of mjk_coulomb_exchange: more than 16 orbital sets, a density
    of the wrong length; the context stays usable afterwards."""
    from openmolcas_amd import MjkError
    n, M = 30, 5
    L = chol_vectors_np(n, M, seed=16)
    jk = JKContext(n, M, device=gpu)
    jk.upload_packed(L)
    C = torch.from_numpy(orbitals_np(n, 3, seed=17))
    dlt = random_density(n, seed=18)
    with pytest.raises(MjkError):
        jk.coulomb_exchange(torch.from_numpy(dlt), [C] * 17)
    with pytest.raises(ValueError):
        jk.coulomb_exchange(torch.from_numpy(dlt[:-1]), [C])
    check(jk, L, dlt, [C.numpy()], n)
    jk.close()
