"""GPU parity of the HIP J/K build against the CPU oracle (FP64).

Every call goes through the C ABI (libmolcas_jk.so).  Tolerances: the FP64
kernels sum in a different order than the oracle, so results agree to rounding:
relative max-abs 1e-11 at oracle sizes; absolute max-abs 1e-8 at full size
(the FP64 target of the J/K build).
"""
import numpy as np
import pytest
import torch

import oracle
from openmolcas_amd import MjkError
from openmolcas_amd.fock import JKContext
from openmolcas_amd.synthetic import (chol_vectors_np, chol_vectors_torch, fold, hcore_np, npair,
                                      orbitals_np, orbitals_torch)

pytestmark = pytest.mark.gpu

RTOL = 1e-11


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return np.max(np.abs(a - b)) / max(1.0, np.max(np.abs(b)))


def make_ctx(L, gpu):
    M, npr = L.shape
    n = int(round((np.sqrt(8 * npr + 1) - 1) / 2))
    ctx = JKContext(n, M, device=gpu)
    ctx.upload_packed(L)
    return ctx


def oracle_J(L, dlt):
    """J[D] from the oracle: cho_fscf with ExFac = 0 and no orbitals."""
    n = int(round((np.sqrt(8 * L.shape[1] + 1) - 1) / 2))
    st = oracle.CholStore.full(L)
    return oracle.cho_fscf(st, np.zeros((1, L.shape[1])), [0], [np.zeros((n, 0))], dlt, 0.0)[0]


def oracle_K(L, C):
    """K = sum_J X X^T from the oracle: cho_fscf with D = 0, ExFac = -1."""
    st = oracle.CholStore.full(L)
    return oracle.cho_fscf(st, np.zeros((1, L.shape[1])), [C.shape[1]], [C],
                           np.zeros(L.shape[1]), -1.0)[0]


# ------------------------------------------------------------------ unit kernels
def test_exchange_single_mfma_tile(gpu):
    # n = 16, one vector, 16 orbitals: one 16x16x4 MFMA tile per k-step; a
    # wrong lane layout of the MFMA operands or result shows here first
    n, M, o = 16, 1, 16
    L = chol_vectors_np(n, M, seed=100)
    C = orbitals_np(n, o, seed=101)
    ctx = make_ctx(L, gpu)
    K = ctx.exchange([torch.from_numpy(C)]).cpu().numpy()[0]
    assert rel(K, oracle_K(L, C)) < RTOL


@pytest.mark.parametrize("n,M,nd", [(5, 3, 1), (37, 50, 2), (130, 41, 3), (200, 64, 9), (300, 130, 5),
                                    (33, 200, 16)])
def test_coulomb_matches_oracle(gpu, n, M, nd):
    L = chol_vectors_np(n, M, seed=n + M)
    rng = np.random.default_rng(n)
    dl = []
    for _ in range(nd):
        C = orbitals_np(n, max(1, n // 4), seed=int(rng.integers(1 << 30)))
        dl.append(fold(2.0 * C @ C.T))
    ctx = make_ctx(L, gpu)
    J, V = ctx.coulomb(torch.from_numpy(np.stack(dl)), want_v=True)
    J = J.cpu().numpy()
    for d in range(nd):
        assert rel(J[d], oracle_J(L, dl[d])) < RTOL
        assert np.allclose(V[:, d].cpu().numpy(), L @ dl[d], rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("n,M,occs", [
    (20, 7, [3]),
    (64, 33, [16, 1]),
    (65, 19, [5, 17, 4]),
    (129, 40, [33]),
    (200, 25, [64, 13, 7, 2]),
    (300, 12, [150]),
    (150, 21, [70]),        # 80-column bank
    (140, 15, [40, 36]),    # 80-column bank, two sets
    (170, 17, [100]),       # 112-column bank
])
def test_exchange_matches_oracle(gpu, n, M, occs):
    L = chol_vectors_np(n, M, seed=7 * n + M)
    Cs = [orbitals_np(n, o, seed=1000 + i) for i, o in enumerate(occs)]
    ctx = make_ctx(L, gpu)
    K = ctx.exchange([torch.from_numpy(c) for c in Cs]).cpu().numpy()
    for d, C in enumerate(Cs):
        assert rel(K[d], oracle_K(L, C)) < RTOL


def test_exchange_chunked_equals_single_pass(gpu):
    n, M = 96, 70
    L = chol_vectors_np(n, M, seed=5)
    C = orbitals_np(n, 20, seed=6)
    ctx = make_ctx(L, gpu)
    K1 = ctx.exchange([torch.from_numpy(C)]).cpu().numpy()
    ctx.set_x_budget(8 * 32 * 128 * 9)  # 9 vectors per chunk
    K2 = ctx.exchange([torch.from_numpy(C)]).cpu().numpy()
    assert rel(K2, K1) < 1e-13
    assert rel(K2[0], oracle_K(L, C)) < RTOL


def test_energy_matches_oracle(gpu):
    n, M, o = 40, 30, 6
    L = chol_vectors_np(n, M, seed=9)
    C = orbitals_np(n, o, seed=10)
    h = hcore_np(n, seed=11)
    P = 2.0 * C @ C.T
    dlt = fold(P)
    st = oracle.CholStore.full(L)
    g = oracle.cho_fscf(st, np.zeros((1, npair(n))), [o], [np.sqrt(2) * C], dlt, 0.5)[0]
    e_ref = oracle.energy(h, dlt, g)
    ctx = make_ctx(L, gpu)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(gpu)
    e = ctx.energy(dev(h), dev(dlt), dev(g)).item()
    assert e == pytest.approx(e_ref, rel=1e-13, abs=1e-12)


# ------------------------------------------------------------------ drop-ins
@pytest.mark.parametrize("nden", [1, 2])
def test_cho_fscf_dropin(gpu, nden):
    n, M = 77, 45
    L = chol_vectors_np(n, M, seed=21)
    Cs = [orbitals_np(n, 9 + 4 * d, seed=22 + d) for d in range(nden)]
    dlt = fold(2.0 * Cs[0] @ Cs[0].T)
    flt0 = np.random.default_rng(23).standard_normal((nden, npair(n)))
    want = oracle.cho_fscf(oracle.CholStore.full(L), flt0, [c.shape[1] for c in Cs], Cs, dlt, 0.5)
    ctx = make_ctx(L, gpu)
    flt = flt0.copy()
    ctx.cho_fscf(nden, flt, [0] * nden, [c.shape[1] for c in Cs], Cs, dlt, 0.5)
    assert rel(flt, want) < RTOL


def test_cho_fscf_dropin_multiple_reduced_sets(gpu):
    rng = np.random.default_rng(30)
    n, M1, M2 = 50, 20, 13
    npr = npair(n)
    s0 = np.arange(npr)
    s1 = np.sort(rng.choice(s0, size=600, replace=False))
    s2 = np.sort(rng.choice(s1, size=150, replace=False))
    L0 = chol_vectors_np(n, M1, seed=31)
    L1 = rng.standard_normal((M2, len(s1))) * 0.05
    L2 = rng.standard_normal((4, len(s2))) * 0.01
    vecs = [(0, v) for v in L0] + [(1, v) for v in L1] + [(2, v) for v in L2]
    st = oracle.CholStore(n, [s0, s1, s2], vecs)
    C = [orbitals_np(n, 11, seed=32)]
    dlt = fold(2.0 * C[0] @ C[0].T)
    flt0 = np.zeros((1, npr))
    want = oracle.cho_fscf(st, flt0, [11], C, dlt, 1.0, batch=8)
    ctx = JKContext(n, M1 + M2 + 4, device=gpu)
    for r, s in enumerate((s0, s1, s2)):
        ctx.set_reduced_set(r, s)
    ctx.upload_vectors(0, 0, L0)
    ctx.upload_vectors(1, M1, L1)
    ctx.upload_vectors(2, M1 + M2, torch.from_numpy(L2).to(gpu))  # device-side upload
    flt = flt0.copy()
    ctx.cho_fscf(1, flt, [0], [11], C, dlt, 1.0)
    assert rel(flt, want) < RTOL


@pytest.mark.parametrize("nden,shells", [(1, None), (2, [5, 9, 3, 20, 13]), (2, [1] * 50)])
def test_cho_lk_scf_dropin(gpu, nden, shells):
    """CHO_LK_SCF without screening: FLT and the shell-pair-blocked KLT."""
    rng = np.random.default_rng(40 + nden)
    n, M1, M2 = 50, 24, 9
    npr = npair(n)
    s0 = np.arange(npr)
    s1 = np.sort(rng.choice(s0, size=500, replace=False))
    L0 = chol_vectors_np(n, M1, seed=41)
    L1 = rng.standard_normal((M2, len(s1))) * 0.05
    st = oracle.CholStore(n, [s0, s1], [(0, v) for v in L0] + [(1, v) for v in L1])
    C = [orbitals_np(n, 9 + 4 * d, seed=42 + d) for d in range(nden)]
    plt = fold(sum(c @ c.T for c in C))
    flt0 = rng.standard_normal((nden, npr))
    want_f, want_k = oracle.cho_lk_scf(st, flt0, [c.shape[1] for c in C], C, plt, 0.5,
                                       nbassh=shells, batch=7)
    ctx = JKContext(n, M1 + M2, device=gpu)
    ctx.set_reduced_set(0, s0)
    ctx.set_reduced_set(1, s1)
    ctx.upload_vectors(0, 0, L0)
    ctx.upload_vectors(1, M1, L1)
    flt = flt0.copy()
    klt = np.zeros((nden, npr))
    ctx.cho_lk_scf(nden, flt, klt, [0] * nden, [c.shape[1] for c in C], C, plt, 0.5,
                   nbassh=shells)
    assert rel(flt, want_f) < RTOL
    assert rel(klt, want_k) < RTOL
    with pytest.raises(MjkError):
        ctx.cho_lk_scf(3, flt, klt, [0] * 3, [1] * 3, C * 3, plt, 0.5)
    with pytest.raises(MjkError):
        ctx.cho_lk_scf(1, flt[:1].copy(), klt[:1].copy(), [0], [3], C[:1], plt, 0.5,
                       nbassh=[10, 10])  # shells do not cover nbas
    ctx.close()


def test_cho_fock_dft_red_dropin(gpu):
    """Coulomb-only drop-in over two reduced sets; FLT is overwritten."""
    rng = np.random.default_rng(60)
    n, M1, M2 = 45, 14, 11
    npr = npair(n)
    s0 = np.sort(rng.choice(np.arange(npr), size=400, replace=False))
    s1 = np.arange(npr)
    L0 = rng.standard_normal((M1, len(s0))) * 0.2
    L1 = chol_vectors_np(n, M2, seed=61)
    st = oracle.CholStore(n, [s0, s1], [(0, v) for v in L0] + [(1, v) for v in L1])
    C = orbitals_np(n, 6, seed=62)
    dlt = fold(2.0 * C @ C.T)
    flt0 = rng.standard_normal(npr)
    want = oracle.cho_fock_dft_red(st, dlt, flt0, batch=5)
    ctx = JKContext(n, M1 + M2, device=gpu)
    ctx.set_reduced_set(0, s0)
    ctx.set_reduced_set(1, s1)
    ctx.upload_vectors(0, 0, L0)
    ctx.upload_vectors(1, M1, L1)
    flt = flt0.copy()
    ctx.cho_fock_dft_red(dlt, flt)
    assert rel(flt, want) < RTOL
    ctx.close()


@pytest.mark.parametrize("exfac,do_active", [(1.0, True), (0.7, True), (0.7, False)])
def test_cho_lk_casscf_dropin(gpu, exfac, do_active):
    n, M = 56, 40
    nfi, na = 9, 5
    L = chol_vectors_np(n, M, seed=70)
    rng = np.random.default_rng(71)
    cmo, _ = np.linalg.qr(rng.standard_normal((n, n)))
    Ci, Ca = cmo[:, :nfi], cmo[:, nfi:nfi + na]
    occ = rng.uniform(0.1, 1.9, na)
    DI, DA = 2.0 * Ci @ Ci.T, (Ca * occ) @ Ca.T
    msq = np.sqrt(2.0) * Ci                    # DECO vectors of DI
    ash2 = Ca * np.sqrt(occ)                   # DECO vectors of DA
    dlt = np.stack([fold(DI), fold(DA)])
    flt0 = rng.standard_normal((2, npair(n)))
    nfro, norb = 1, n - 1
    want_f, want_w = oracle.cho_lk_casscf(oracle.CholStore.full(L), dlt, flt0, msq, -0.5, nfi, na, na,
                                          Ca, ash2, do_active, cmo, nfro, norb, exfac, batch=9)
    ctx = make_ctx(L, gpu)
    flt = flt0.copy()
    w = np.zeros(norb * na * (na * (na + 1) // 2))
    ctx.cho_lk_casscf(dlt, flt, msq, w, -0.5, nfi, na, na, Ca, ash2, do_active, cmo, nfro, norb,
                      exfac=exfac)
    assert rel(flt, want_f) < RTOL
    assert rel(w, want_w) < RTOL
    ctx.close()


def test_cho_fmcscf_dropin_with_puvx(gpu):
    n, M = 60, 35
    nk, na = 7, 5
    L = chol_vectors_np(n, M, seed=50)
    rng = np.random.default_rng(51)
    cmo, _ = np.linalg.qr(rng.standard_normal((n, n)))
    Ci, Ca = cmo[:, :nk], cmo[:, nk:nk + na]
    occ = rng.uniform(0.1, 1.9, na)
    DI, DA = 2.0 * Ci @ Ci.T, (Ca * occ) @ Ca.T
    Cm = Ca * np.sqrt(occ)
    dlt = np.stack([fold(DI), fold(DA)])
    flt0 = rng.standard_normal((2, npair(n)))
    nfro, norb = 2, n - 3
    want_f, want_w = oracle.cho_fmcscf(oracle.CholStore.full(L), flt0, nk, na, -0.5, dlt, True,
                                       np.sqrt(2.0) * Ci, na, Cm, Ca, cmo, nfro, norb, batch=9)
    ctx = make_ctx(L, gpu)
    flt = flt0.copy()
    w = np.zeros(norb * na * na * (na + 1) // 2)
    ctx.cho_fmcscf(flt, 0, nk, na, -0.5, dlt, True, np.sqrt(2.0) * Ci, na, Cm, Ca, w, cmo, nfro, norb)
    assert rel(flt, want_f) < RTOL
    assert rel(w, want_w) < RTOL


# ------------------------------------------------------------------ error paths
def test_error_codes(gpu):
    n, M = 10, 6
    L = chol_vectors_np(n, M, seed=1)
    ctx = JKContext(n, M, device=gpu)
    ctx.set_reduced_set(0, np.arange(npair(n)))
    ctx.upload_vectors(0, 0, L[:3])  # short: 3 of 6 vectors
    with pytest.raises(MjkError) as e:
        ctx.coulomb(torch.zeros(npair(n), dtype=torch.float64))
    assert e.value.rc == 77
    ctx.upload_vectors(0, 3, L[3:])
    flt = np.zeros((3, npair(n)))
    with pytest.raises(MjkError) as e:  # nDen not in {1,2}: the reference abends
        ctx.cho_fscf(3, flt, [0, 0, 0], [1, 1, 1], [np.zeros((n, 1))] * 3, np.zeros(npair(n)), 1.0)
    assert e.value.rc == 1
    with pytest.raises(MjkError) as e:  # vector range beyond capacity
        ctx.upload_vectors(0, 5, L[:3])
    assert e.value.rc == 1
    with pytest.raises(MjkError) as e:  # undeclared reduced set
        ctx.upload_vectors(4, 0, L[:1])
    assert e.value.rc == 3
    with pytest.raises(MjkError) as e:  # LT index out of range
        ctx.set_reduced_set(1, np.array([0, npair(n)]))
    assert e.value.rc == 1


# ------------------------------------------------------------------ full size
def _torch_jk(L, dlt, C, n, chunk=64):
    """J and K in torch FP64 (rocBLAS), from the unpacked vectors."""
    V = L @ dlt
    J = V @ L
    ii, jj = torch.tril_indices(n, n, device=L.device)
    K = torch.zeros((n, n), dtype=torch.float64, device=L.device)
    for j0 in range(0, L.shape[0], chunk):
        blk = L[j0:j0 + chunk]
        Lf = torch.zeros((blk.shape[0], n, n), dtype=torch.float64, device=L.device)
        Lf[:, ii, jj] = blk
        Lf[:, jj, ii] = blk
        X = torch.matmul(Lf, C.contiguous())  # (b, n, o)
        Xt = X.transpose(0, 1).reshape(n, -1).contiguous()  # (n, b*o)
        K += Xt @ Xt.t().contiguous()
    return J, V, K[ii, jj]


def test_full_size_n512_m2048_o64(gpu):
    """Synthetic single FP64 J/K build at n=512, nchol=2048, nocc=64 vs torch FP64."""
    n, M, o = 512, 2048, 64
    L = chol_vectors_torch(n, M, seed=2024, device=gpu)
    C = orbitals_torch(n, o, seed=7, device=gpu)
    P = C @ C.T
    ii, jj = torch.tril_indices(n, n, device=gpu)
    dlt = P[ii, jj] * torch.where(ii == jj, 1.0, 2.0).to(torch.float64)
    ctx = JKContext(n, M, device=gpu)
    ctx.upload_packed(L)
    J, V = ctx.coulomb(dlt.unsqueeze(0), want_v=True)
    K = ctx.exchange([C])
    Jr, Vr, Kr = _torch_jk(L, dlt, C, n)
    assert torch.max(torch.abs(J[0] - Jr)).item() <= 1e-8
    assert torch.max(torch.abs(K[0] - Kr)).item() <= 1e-8
    # size-independent identities: tr(D J[D]) = sum_J V_J^2 and tr(D K) >= 0
    assert torch.dot(dlt, J[0]).item() == pytest.approx(torch.sum(V[:, 0] ** 2).item(), rel=1e-12)
    assert torch.dot(dlt, K[0]).item() > 0.0
