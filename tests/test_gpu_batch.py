"""Batched builds: the state-averaged density batch and the uneven molecule batch.

* State-averaged CASSCF-style batch: nbas=800, nchol=3600, one inactive density
  of rank 120 and 8 active densities of ranks U{4..16} with occupations in
  (0,2); all 9 J and 9 K in one pass over the vectors (one Coulomb pass, one
  half-transform, one SYRK launch).  FP64 checked against torch FP64 (rocBLAS)
  to max-abs 1e-8; FP32 exchange to relative 2e-6; the precision ladder tuned
  per density keeps every density's exchange energy within 1e-6 Eh of FP64.
* Uneven molecule batch: independent systems (nbas U[100,600], nchol U[4,6]
  nbas, nocc 10-20% nbas) sharing one workspace; forward + backward captured
  in one HIP graph; every replay must reproduce the eager step bit for bit.
  At full size (64 systems) every system's FP64 energy and dE/dC are checked
  against an independent torch FP64 build of J and K, and the tuned ladder
  keeps every system within 1e-6 Eh (dE/dC within 1e-5, dE/dL of four systems
  within 1e-5 relative, its tail vectors from the FP32 backward) of FP64.
"""
import numpy as np
import pytest
import torch

from openmolcas_amd import ladder
from openmolcas_amd.batch import MoleculeBatch, uneven_shapes
from openmolcas_amd.fock import JKContext
from openmolcas_amd.synthetic import chol_vectors_torch, npair

pytestmark = pytest.mark.gpu


def _fold(P, n):
    ii, jj = torch.tril_indices(n, n, device=P.device)
    return (P[ii, jj] * torch.where(ii == jj, 1.0, 2.0).to(torch.float64)).contiguous()


def _torch_k(L, Cs, n, chunk=48):
    ii, jj = torch.tril_indices(n, n, device=L.device)
    Ks = [torch.zeros((n, n), dtype=torch.float64, device=L.device) for _ in Cs]
    for j0 in range(0, L.shape[0], chunk):
        blk = L[j0:j0 + chunk]
        Lf = torch.zeros((blk.shape[0], n, n), dtype=torch.float64, device=L.device)
        Lf[:, ii, jj] = blk
        Lf[:, jj, ii] = blk
        for K, C in zip(Ks, Cs):
            X = torch.matmul(Lf, C.contiguous())
            Xt = X.transpose(0, 1).reshape(n, -1).contiguous()
            K += Xt @ Xt.t().contiguous()
    return [K[ii, jj] for K in Ks]


def test_state_averaged_batch_n800_m3600(gpu):
    n, M = 800, 3600
    rng = np.random.default_rng(11)
    ranks = [120] + [int(r) for r in rng.integers(4, 17, size=8)]
    L = chol_vectors_torch(n, M, seed=12, device=gpu)
    g = torch.Generator(device=gpu)
    g.manual_seed(13)
    Q, _ = torch.linalg.qr(torch.randn((n, sum(ranks)), generator=g, dtype=torch.float64, device=gpu))
    Cs, dl, o0 = [], [], 0
    for d, r in enumerate(ranks):
        occ = torch.full((r,), 2.0, dtype=torch.float64, device=gpu) if d == 0 else \
            torch.tensor(rng.uniform(0.05, 1.95, r), dtype=torch.float64, device=gpu)
        Cd = (Q[:, o0:o0 + r] * torch.sqrt(occ)).contiguous()  # D_d = C_d C_d^T
        o0 += r
        Cs.append(Cd)
        dl.append(_fold(Cd @ Cd.T, n))
    jk = JKContext(n, M, device=gpu)
    jk.upload_packed(L)
    jk.set_profiling(True)
    D = torch.stack(dl)
    J = jk.coulomb(D)
    K = jk.exchange(Cs)
    times = jk.phase_times()
    jk.set_profiling(False)
    assert times["half_transform"][1] == 1 and times["exchange"][1] == 1  # one pass, one SYRK launch
    Jr = (L @ D.T).T @ L
    Kr = _torch_k(L, Cs, n)
    assert torch.max(torch.abs(J - Jr)).item() <= 1e-8
    for d in range(len(ranks)):
        assert torch.max(torch.abs(K[d] - Kr[d])).item() <= 1e-8, d
    K32 = jk.exchange(Cs, precision="fp32")
    for d in range(len(ranks)):
        err = torch.max(torch.abs(K32[d] - Kr[d])).item() / torch.max(torch.abs(Kr[d])).item()
        assert err < 2e-6, (d, err)
    # precision ladder, one split for all nine densities, verified per density:
    # dE_d = 1/2 |sum_LT (K_d - K64_d) D_d| (the closed-shell exchange energy of D_d)
    plan = ladder.tune(jk, Cs, D, factor=0.5, tol=1e-6, per_density=True)
    assert plan.j_fp32 < M
    Kl = jk.exchange_mixed(Cs, plan.j_fp32, plan.j_bf16)
    for d in range(len(ranks)):
        de = 0.5 * abs(torch.sum((Kl[d] - K[d]) * D[d]).item())
        assert de <= 1e-6, (d, de)
    print(f"state-averaged ladder: split {plan.j_fp32} / {plan.j_bf16} of {M}, "
          f"max per-density |dE| {plan.measured:.2e}")


def _run_batch_checks(batch, precision):
    batch.step(precision)
    batch.stream.synchronize()
    E0 = batch.E.clone()
    dC0 = [d.clone() for d in batch.dC]
    # dL is as large as the vectors: compare an exact integer checksum of its bits
    dL0 = [d.view(torch.int64).sum().item() if d is not None else None for d in batch.dL]
    batch.capture(precision)
    for rep in range(3):
        for d in batch.dC:  # eager work between replays
            d.zero_()
        batch.E.zero_()
        torch.cuda.synchronize()
        batch.replay()
        torch.cuda.synchronize()
        assert torch.equal(batch.E, E0), rep
        for a, b in zip(batch.dC, dC0):
            assert torch.equal(a, b), rep
        for a, b in zip(batch.dL, dL0):
            if a is not None:
                assert a.view(torch.int64).sum().item() == b, rep
    return E0, dC0


def test_molecule_batch_graph_small(gpu):
    shapes = [(40, 170, 6), (77, 330, 12), (130, 600, 20), (64, 300, 9)]
    batch = MoleculeBatch(shapes, device=gpu, seed=21)
    E0, dC0 = _run_batch_checks(batch, "fp64")
    # each system against an independent context (not shared, not captured)
    for i, (n, M, o) in enumerate(shapes):
        jk = JKContext(n, M, device=gpu)
        jk.upload_packed(batch.L[i].clone())
        C, h = batch.C[i], batch.h[i]
        dlt = jk.density(C, 2.0)
        J = jk.coulomb(dlt)
        K = jk.exchange([C])
        G = torch.zeros_like(J)
        jk.fock_update(G, J, [0], [1.0], K, [-1.0])
        e = jk.energy(h, dlt, G).item()
        assert e == pytest.approx(E0[i].item(), rel=1e-13, abs=1e-10)
        dc = jk.grad_orbitals(C, h + G, 4.0)
        assert torch.max(torch.abs(dc - dC0[i])).item() <= 1e-10 * max(1.0, torch.max(torch.abs(dc)).item())
        jk.close()
    batch.close()


def _torch_system(L, C, h, n):
    """E and dE/dC of one closed-shell system from torch FP64 alone: P = 2 C C^T,
    J = sum_J L^J (L^J . fold(P)), K = K[C], G = J - K, E = h.D + 1/2 G.D
    (D = fold(P)), dE/dC = 4 (h + G) C."""
    ii, jj = torch.tril_indices(n, n, device=L.device)
    D = _fold(2.0 * C @ C.T, n)
    J = (L @ D) @ L
    K = _torch_k(L, [C], n, chunk=128)[0]
    G = J - K
    E = torch.dot(h.reshape(-1), D) + 0.5 * torch.dot(G, D)
    F = torch.zeros((n, n), dtype=torch.float64, device=L.device)
    F[ii, jj] = h.reshape(-1) + G
    F[jj, ii] = h.reshape(-1) + G
    return E.item(), 4.0 * F @ C


def test_uneven_molecule_batch_64_graph(gpu):
    """64 systems, nbas U[100,600], forward + backward in one captured graph:
    FP64 (each system also against torch FP64), FP32 exchange, and the tuned
    precision ladder at |dE| <= 1e-6 Eh per system."""
    shapes = uneven_shapes(64, seed=5)
    batch = MoleculeBatch(shapes, device=gpu, seed=3)
    batch.autotune("fp64")
    E64, dC64 = _run_batch_checks(batch, "fp64")
    worst_e, worst_c = 0.0, 0.0
    for i, (n, M, o) in enumerate(shapes):
        e, dc = _torch_system(batch.L[i], batch.C[i], batch.h[i], n)
        worst_e = max(worst_e, abs(e - E64[i].item()) / max(1.0, abs(e)))
        worst_c = max(worst_c, torch.max(torch.abs(dc - dC64[i])).item() / torch.max(torch.abs(dc)).item())
    print(f"64-system batch against torch FP64: max rel |dE| {worst_e:.2e}, max rel dE/dC {worst_c:.2e}")
    assert worst_e <= 1e-11 and worst_c <= 1e-10
    # dE/dL of four systems (smallest, largest, two between) kept for the ladder
    order = sorted(range(len(shapes)), key=lambda i: shapes[i][0])
    keep = [order[0], order[len(order) // 3], order[2 * len(order) // 3], order[-1]]
    dL64 = {i: batch.dL[i].clone() for i in keep}
    E32, dC32 = _run_batch_checks(batch, "fp32")
    dE = torch.max(torch.abs(E32 - E64)).item()
    dCerr = max(torch.max(torch.abs(a - b)).item() for a, b in zip(dC32, dC64))
    print(f"64-system batch: max |dE| fp32 vs fp64 = {dE:.3e}, max |d(dE/dC)| = {dCerr:.3e}")
    # FP32 throughout misses 1e-6 Eh (that is what the ladder is for); its error
    # stays at the FP32 rounding level of the exchange energy
    assert dE <= 1e-6 * torch.max(torch.abs(E64)).item()
    batch.tune_ladder(tol=1e-6)
    El, dCl = _run_batch_checks(batch, "ladder")
    dE = torch.max(torch.abs(El - E64)).item()
    dCerr = max(torch.max(torch.abs(a - b)).item() for a, b in zip(dCl, dC64))
    # dE/dL past each ladder's FP64 head comes from the FP32 backward products
    dLerr = max((torch.max(torch.abs(batch.dL[i] - dL64[i])) / torch.max(torch.abs(dL64[i]))).item()
                for i in keep)
    print(f"64-system batch: max |dE| ladder vs fp64 = {dE:.3e}, max |d(dE/dC)| = {dCerr:.3e}, "
          f"max rel d(dE/dL) = {dLerr:.3e}")
    assert dE <= 1e-6 and dCerr <= 1e-5 and dLerr <= 1e-5
    batch.close()
