"""Backward of the closed-shell energy (hand-written HIP gradient kernels).

Checked against central finite differences of the oracle energy (step 1e-5)
and against the analytic orbital gradient 4 F C.  Tolerances are written per
check: FD error of a step-1e-5 central difference is O(eps^2 * E''') plus
O(roundoff / eps), so |fd - analytic| <= 1e-6 + 1e-6 |analytic| at oracle
sizes; at full size (n=1200, nchol=6000, nocc=150) FD is taken on the GPU's
FP64 energy and the bound is 2e-4 + 1e-6 |analytic| (E ~ 5e6 Eh, so roundoff
of E / step ~ 1e-4).
"""
import numpy as np
import pytest
import torch

import oracle
from openmolcas_amd.autograd import closed_shell_energy
from openmolcas_amd.fock import JKContext
from openmolcas_amd.synthetic import (chol_vectors_np, chol_vectors_torch, fold, hcore_np, npair,
                                      orbitals_np, orbitals_torch, pack, unpack)

pytestmark = pytest.mark.gpu


def oracle_energy(C, L, h):
    n = C.shape[0]
    P = 2.0 * C @ C.T
    st = oracle.CholStore.full(L)
    g = oracle.cho_fscf(st, np.zeros((1, npair(n))), [C.shape[1]], [np.sqrt(2.0) * C], fold(P), 0.5)[0]
    return oracle.energy(h, fold(P), g), g


def test_energy_gradients_match_finite_differences(gpu):
    n, M, o = 24, 20, 4
    L = chol_vectors_np(n, M, seed=70)
    C = orbitals_np(n, o, seed=71)
    h = hcore_np(n, seed=72)
    jk = JKContext(n, M, device=gpu)
    Ct = torch.tensor(C, device=gpu, requires_grad=True)
    Lt = torch.tensor(L, device=gpu, requires_grad=True)
    ht = torch.tensor(h, device=gpu, requires_grad=True)
    E = closed_shell_energy(Ct, Lt, ht, jk)
    E.backward()
    e_ref, g_ref = oracle_energy(C, L, h)
    assert E.item() == pytest.approx(e_ref, rel=1e-12, abs=1e-10)
    # dE/dC = 4 F C (analytic) with F = h + G from the oracle
    F = unpack(h + g_ref, n)
    assert np.allclose(Ct.grad.cpu().numpy(), 4.0 * F @ C, rtol=1e-10, atol=1e-10)
    # dE/dh = fold(P)
    assert np.allclose(ht.grad.cpu().numpy(), fold(2.0 * C @ C.T), rtol=1e-12, atol=1e-13)
    eps = 1e-5
    rng = np.random.default_rng(73)
    dC = Ct.grad.cpu().numpy()
    for _ in range(8):
        a, k = int(rng.integers(n)), int(rng.integers(o))
        Cp, Cm = C.copy(), C.copy()
        Cp[a, k] += eps
        Cm[a, k] -= eps
        fd = (oracle_energy(Cp, L, h)[0] - oracle_energy(Cm, L, h)[0]) / (2 * eps)
        assert abs(fd - dC[a, k]) <= 1e-6 + 1e-6 * abs(dC[a, k])
    dL = Lt.grad.cpu().numpy()
    diag = np.array([p * (p + 1) // 2 + p for p in range(n)])
    picks = [(int(rng.integers(M)), int(rng.integers(npair(n)))) for _ in range(10)]
    picks += [(int(rng.integers(M)), int(diag[rng.integers(n)])) for _ in range(4)]
    for j, p in picks:
        Lp, Lm = L.copy(), L.copy()
        Lp[j, p] += eps
        Lm[j, p] -= eps
        fd = (oracle_energy(C, Lp, h)[0] - oracle_energy(C, Lm, h)[0]) / (2 * eps)
        assert abs(fd - dL[j, p]) <= 1e-6 + 1e-6 * abs(dL[j, p]), (j, p, fd, dL[j, p])


def test_vector_gradient_chunked_equals_single_pass(gpu):
    n, M, o = 70, 37, 9
    L = torch.tensor(chol_vectors_np(n, M, seed=80), device=gpu)
    C = torch.tensor(orbitals_np(n, o, seed=81), device=gpu)
    jk = JKContext(n, M, device=gpu)
    jk.upload_packed(L)
    dlt = jk.density(C, 2.0)
    _, V = jk.coulomb(dlt, want_v=True)
    g1 = jk.grad_vectors(C, V[:, 0].contiguous(), 1.0, -2.0).cpu().numpy()
    jk.set_x_budget(8 * 6 * (2 * 16 * 128 + 16 * 16))  # 6 vectors per chunk
    g2 = jk.grad_vectors(C, V[:, 0].contiguous(), 1.0, -2.0).cpu().numpy()
    assert np.max(np.abs(g1 - g2)) <= 1e-13 * max(1.0, np.max(np.abs(g1)))


@pytest.mark.parametrize("fused", [True, False])
def test_saved_z_gradient_equals_recompute(gpu, fused):
    """Z^J = C^T L^J C kept by the forward (mjk_coulomb_exchange_z or
    mjk_exchange_z) against numpy, K unchanged by the Z output, and the
    vector gradient from Z (mjk_grad_vectors_z) equal to the one that
    recomputes X^J (the same half-transform and GEMMs, so within 1e-13)."""
    n, M, o = 150, 90, 21
    Ln = chol_vectors_np(n, M, seed=83)
    Cn = orbitals_np(n, o, seed=84)
    L = torch.tensor(Ln, device=gpu)
    C = torch.tensor(Cn, device=gpu)
    jk = JKContext(n, M, device=gpu)
    jk.upload_packed(L)
    jk.set_x_budget(8 * 16 * (2 * 32 * 256 + 32 * 32))  # several chunks on both sides
    dlt = jk.density(C, 2.0)
    Z = torch.zeros((M, o, o), dtype=torch.float64, device=gpu)
    if fused:
        J, K, V = jk.coulomb_exchange(dlt, [C], want_v=True, z=Z)
        J0, K0, V0 = jk.coulomb_exchange(dlt, [C], want_v=True)
        assert torch.equal(J, J0) and torch.equal(V, V0)
    else:
        _, V = jk.coulomb(dlt, want_v=True)
        K = jk.exchange([C], z=Z)
        K0 = jk.exchange([C])
    assert torch.equal(K, K0)
    Zr = np.einsum("ak,jab,bl->jkl", Cn, np.stack([unpack(Ln[j], n) for j in range(M)]), Cn)
    assert np.max(np.abs(Z.cpu().numpy() - Zr)) <= 1e-12 * np.max(np.abs(Zr))
    v = V[:, 0].contiguous()
    g_rec = jk.grad_vectors(C, v, 1.0, -2.0)
    g_z = jk.grad_vectors(C, v, 1.0, -2.0, z=Z)
    assert torch.max(torch.abs(g_z - g_rec)).item() <= 1e-13 * torch.max(torch.abs(g_rec)).item()
    # a reduced-precision forward keeps Z at that precision (FP32 X, FP64 product)
    Z32 = torch.zeros_like(Z)
    jk.exchange([C], precision="fp32", z=Z32)
    assert torch.max(torch.abs(Z32 - Z)).item() <= 1e-5 * torch.max(torch.abs(Z)).item()
    with pytest.raises(ValueError):
        jk.exchange([C], z=torch.zeros((M, o, o + 1), dtype=torch.float64, device=gpu))
    jk.close()


@pytest.mark.parametrize("n,M,o", [(40, 7, 5), (64, 5, 16), (130, 6, 33), (200, 5, 64), (150, 4, 70),
                                   (333, 3, 100), (170, 3, 112), (129, 3, 128), (257, 2, 129)])
def test_saved_z_matches_dense(gpu, n, M, o):
    """Z^J = C^T L^J C written by the forward (lower 16 x 16 blocks mirrored,
    every block-row count 1..8; nocc 129 takes the general GEMM) against
    numpy, FP64 to 1e-12 and from the FP32 half-transform to 1e-5."""
    Ln = chol_vectors_np(n, M, seed=90 + o)
    Cn = orbitals_np(n, o, seed=91 + o)
    jk = JKContext(n, M, device=gpu)
    jk.upload_packed(torch.tensor(Ln, device=gpu))
    C = torch.tensor(Cn, device=gpu)
    Zr = np.einsum("ak,jab,bl->jkl", Cn, np.stack([unpack(Ln[j], n) for j in range(M)]), Cn)
    for prec, tol in (("fp64", 1e-12), ("fp32", 1e-5)):
        Z = torch.full((M, o, o), float("nan"), dtype=torch.float64, device=gpu)
        jk.exchange([C], precision=prec, z=Z)
        err = np.max(np.abs(Z.cpu().numpy() - Zr)) / np.max(np.abs(Zr))
        assert err <= tol, (prec, err)
    jk.close()


@pytest.mark.parametrize("n,M,o", [(1, 2, 1), (64, 3, 16), (70, 9, 5), (130, 7, 33), (200, 6, 64),
                                   (333, 5, 100), (129, 4, 128), (257, 3, 129)])
def test_vector_gradient_from_z_matches_dense(gpu, n, M, o):
    """dE/dL from a given Z (mjk_grad_vectors_z) against the dense formula in
    numpy: G^J = s (alpha V_J P + beta C Z^J C^T), packed lower triangle with
    s = 1 on the diagonal and 2 off it.  The orbital counts cover every width
    of the fused kernel (nocc up to 128, in 16-column steps) and the two-GEMM
    path beyond it (129); bases cover one AO and ragged 64-AO tiles.  Z is an
    arbitrary (non-symmetric) matrix, so a transposed operand would show."""
    rng = np.random.default_rng(n + 7 * o)
    Cn = rng.standard_normal((n, o))
    Zn = rng.standard_normal((M, o, o))
    Pn = pack(rng.standard_normal((n, n)))
    vn = rng.standard_normal(M)
    alpha, beta = 0.75, -2.0
    jk = JKContext(n, M, device=gpu)
    jk.upload_packed(np.zeros((M, npair(n))))
    got = jk.grad_vectors(torch.tensor(Cn, device=gpu), torch.tensor(vn, device=gpu), alpha, beta,
                          p_lt=torch.tensor(Pn, device=gpu), z=torch.tensor(Zn, device=gpu)).cpu().numpy()
    i, j = np.tril_indices(n)
    s = np.where(i == j, 1.0, 2.0)
    want = np.stack([s * (alpha * vn[J] * Pn + beta * (Cn @ Zn[J] @ Cn.T)[i, j]) for J in range(M)])
    assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want))
    jk.close()


@pytest.mark.parametrize("n,M,o,split", [(70, 9, 5, 9), (130, 7, 33, 3), (200, 6, 64, 0), (129, 4, 128, 2),
                                         (257, 3, 96, 1)])
def test_vector_gradient_fp32_tail(gpu, n, M, o, split):
    """mjk_grad_vectors_z_mixed: vectors below the split are bit-identical to
    mjk_grad_vectors_z; from the split on, C Z^J C^T is formed from FP32
    operands with FP32 sums of o terms twice, so each element stays within the
    a-priori bound (2 o + 4) 2^-24 |beta| s (|C| |Z^J| |C|^T) of the dense FP64
    value (the V P term is FP64)."""
    rng = np.random.default_rng(n + 3 * o + split)
    Cn = rng.standard_normal((n, o))
    Zn = rng.standard_normal((M, o, o))
    Pn = pack(rng.standard_normal((n, n)))
    vn = rng.standard_normal(M)
    alpha, beta = 0.75, -2.0
    jk = JKContext(n, M, device=gpu)
    jk.upload_packed(np.zeros((M, npair(n))))
    args = (torch.tensor(Cn, device=gpu), torch.tensor(vn, device=gpu), alpha, beta)
    kw = dict(p_lt=torch.tensor(Pn, device=gpu), z=torch.tensor(Zn, device=gpu))
    full = jk.grad_vectors(*args, **kw).cpu().numpy()
    got = jk.grad_vectors(*args, fp32_from=split, **kw).cpu().numpy()
    assert np.array_equal(got[:split], full[:split])
    i, j = np.tril_indices(n)
    s = np.where(i == j, 1.0, 2.0)
    for J in range(split, M):
        want = s * (alpha * vn[J] * Pn + beta * (Cn @ Zn[J] @ Cn.T)[i, j])
        bound = (2 * o + 4) * 2.0 ** -24 * abs(beta) * s * (np.abs(Cn) @ np.abs(Zn[J]) @ np.abs(Cn).T)[i, j]
        assert np.all(np.abs(got[J] - want) <= bound + 1e-300), J
        if o >= 16:  # and it is an FP32 result, not the FP64 one
            assert np.max(np.abs(got[J] - full[J])) > 0.0
    jk.close()


def test_full_size_backward_n1200_m6000_o150(gpu):
    """Mixed precision plus backward workload: nbas=1200, nchol=6000, nocc=150.
    autograd dE/dC and dE/dL against central differences (step 1e-5) on 32
    random entries each, using the FP64 forward on the GPU."""
    n, M, o = 1200, 6000, 150
    L = chol_vectors_torch(n, M, seed=3, device=gpu)
    C = orbitals_torch(n, o, seed=4, device=gpu)
    g = torch.Generator(device=gpu)
    g.manual_seed(5)
    hs = torch.randn((n, n), generator=g, dtype=torch.float64, device=gpu)
    ii, jj = torch.tril_indices(n, n, device=gpu)
    h = ((hs + hs.T) / (2.0 * n))[ii, jj].contiguous()
    jk = JKContext(n, M, device=gpu)
    Ct = C.clone().requires_grad_(True)
    Lt = L.requires_grad_(True)
    E = closed_shell_energy(Ct, Lt, h, jk)
    E.backward()
    dC = Ct.grad
    dL = Lt.grad
    eps = 1e-5
    rng = np.random.default_rng(6)

    def energy_C(Cm):
        return closed_shell_energy(Cm, Lt.detach(), h, jk).item()

    with torch.no_grad():
        for _ in range(32):
            a, k = int(rng.integers(n)), int(rng.integers(o))
            Cp = C.clone(); Cp[a, k] += eps
            Cm = C.clone(); Cm[a, k] -= eps
            fd = (energy_C(Cp) - energy_C(Cm)) / (2 * eps)
            an = dC[a, k].item()
            assert abs(fd - an) <= 2e-4 + 1e-6 * abs(an), (a, k, fd, an)
        for _ in range(32):
            j, p = int(rng.integers(M)), int(rng.integers(npair(n)))
            row = Lt[j].detach().clone()
            vals = []
            for s in (+1.0, -1.0):
                r = row.clone(); r[p] += s * eps
                jk.upload_vectors(0, j, r.unsqueeze(0))
                vals.append(_energy_resident(jk, C, h))
            jk.upload_vectors(0, j, row.unsqueeze(0))
            fd = (vals[0] - vals[1]) / (2 * eps)
            an = dL[j, p].item()
            assert abs(fd - an) <= 2e-4 + 1e-6 * abs(an), (j, p, fd, an)


def _energy_resident(jk, C, h):
    """Energy with the vectors currently resident in the context (FP64)."""
    dlt = jk.density(C, scale=2.0)
    J = jk.coulomb(dlt)
    K = jk.exchange([C])
    G = torch.zeros_like(J)
    jk.fock_update(G, J, [0], [1.0], K, [-1.0])
    return jk.energy(h.reshape(1, -1), dlt, G).item()
