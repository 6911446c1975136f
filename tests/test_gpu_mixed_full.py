"""Mixed precision at the north-star bound, at full size, against the FP64 GPU
kernel (whose parity with the oracle and with torch FP64 is tested elsewhere):

* n=1200, nchol=6000, nocc=150 (P = 2CC^T): the tuned ladder (FP64 / FP32 /
  split-BF16 segments) keeps |dE| <= 1e-6 Eh and the orbital gradient
  dE/dC = 4FC within 1e-5 max-abs;
* n=800, nchol=3600, the 9 state-averaged densities (rank 120 and 8 ranks
  U{4..16}, occupations in (0,2)) in one exchange call: every density's
  exchange energy -1/4 tr(D_d K_d) within 1e-6 Eh;
* the 64-system uneven batch (graph-captured forward + backward): every
  system's FP64 energy and dE/dC against an independent torch FP64
  computation of J and K at full size, and every system's ladder energy within
  1e-6 Eh of its FP64 energy.
The reference computes the exchange in FP64 only (src/fock_util/
cho_fscf.F90:250,276); the FP64 side of each comparison is that arithmetic."""
import numpy as np
import pytest
import torch

from openmolcas_amd import ladder
from openmolcas_amd.batch import MoleculeBatch, uneven_shapes
from openmolcas_amd.fock import JKContext
from openmolcas_amd.synthetic import chol_vectors_torch, orbitals_torch

pytestmark = pytest.mark.gpu
TOL_E = 1e-6
TOL_DC = 1e-5


def sym_h(n, seed, dev):
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    hs = torch.randn((n, n), generator=g, dtype=torch.float64, device=dev)
    ii, jj = torch.tril_indices(n, n, device=dev)
    return ((hs + hs.T) / (2.0 * n))[ii, jj].reshape(1, -1).contiguous()


def closed_shell(jk, C, h, dlt, J, K):
    """E = h.D + 1/2 (J - K[CC^T]).D with D = fold(2CC^T), and dE/dC = 4 F C."""
    G = torch.zeros_like(J)
    jk.fock_update(G, J, [0], [1.0], K, [-1.0])
    return jk.energy(h, dlt, G).item(), jk.grad_orbitals(C, h + G, 4.0)


def test_ladder_n1200_m6000_o150(gpu):
    n, M, o = 1200, 6000, 150
    jk = JKContext(n, M, device=gpu)
    L = jk.device_vectors()
    chol_vectors_torch(n, M, 3, device=gpu, out=L)
    C = orbitals_torch(n, o, 4, device=gpu)
    h = sym_h(n, 5, gpu)
    dlt = jk.density(C, 2.0)
    J = jk.coulomb(dlt)
    e64, dc64 = closed_shell(jk, C, h, dlt, J, jk.exchange([C]))
    plan = ladder.tune(jk, [C], dlt, factor=0.5, tol=TOL_E)
    assert plan.j_fp32 < M, "the ladder kept every vector in FP64"
    jk.set_ladder(plan.j_fp32, plan.j_bf16)
    e, dc = closed_shell(jk, C, h, dlt, J, jk.exchange([C], precision="ladder"))
    dcerr = torch.max(torch.abs(dc - dc64)).item()
    print(f"n=1200 ladder: j_fp32={plan.j_fp32} j_bf16={plan.j_bf16} |dE|={abs(e - e64):.3e} "
          f"max|d dE/dC|={dcerr:.3e}")
    assert abs(e - e64) <= TOL_E
    assert dcerr <= TOL_DC
    del L
    jk.close()


def test_ladder_state_averaged_9_densities(gpu):
    n, M = 800, 3600
    rng = np.random.default_rng(11)
    ranks = [120] + [int(r) for r in rng.integers(4, 17, size=8)]
    jk = JKContext(n, M, device=gpu)
    L = jk.device_vectors()
    chol_vectors_torch(n, M, 12, device=gpu, out=L)
    Q = orbitals_torch(n, sum(ranks), 13, device=gpu)
    Cs, Ds, o0 = [], [], 0
    for d, r in enumerate(ranks):
        occ = torch.full((r,), 2.0, dtype=torch.float64, device=gpu) if d == 0 else \
            torch.tensor(rng.uniform(0.05, 1.95, r), dtype=torch.float64, device=gpu)
        Cd = (Q[:, o0:o0 + r] * torch.sqrt(occ)).contiguous()
        o0 += r
        Cs.append(Cd)
        Ds.append(jk.density(Cd, 1.0))                    # fold(C_d C_d^T), occupations inside C_d
    K64 = jk.exchange(Cs)
    # each density's own ladder, then the most conservative split for the joint call
    plans = [ladder.tune(jk, [Cs[d]], Ds[d], factor=0.25, tol=TOL_E) for d in range(len(ranks))]
    j32 = max(p.j_fp32 for p in plans)
    jb = max(max(p.j_bf16 for p in plans), j32)
    K = jk.exchange_mixed(Cs, j32, jb)
    for d in range(len(ranks)):
        de = ladder.exchange_energy_error(K[d:d + 1], K64[d:d + 1], Ds[d], 0.25)
        print(f"density {d} rank {ranks[d]}: |dE| = {de:.3e}")
        assert de <= TOL_E, (d, de)
    assert j32 < M, "the ladder kept every vector in FP64"
    del L
    jk.close()


def _torch_jk(Lp, C, n, chunk=32):
    """J[P] and K[CC^T] on the full square from packed vectors, FP64 torch."""
    ii, jj = torch.tril_indices(n, n, device=Lp.device)
    P = 2.0 * C @ C.T
    w = torch.where(ii == jj, 1.0, 2.0).to(torch.float64)
    V = Lp @ (P[ii, jj] * w)
    Jl = V @ Lp
    J = torch.zeros((n, n), dtype=torch.float64, device=Lp.device)
    J[ii, jj] = Jl
    J[jj, ii] = Jl
    K = torch.zeros((n, n), dtype=torch.float64, device=Lp.device)
    for j0 in range(0, Lp.shape[0], chunk):
        blk = Lp[j0:j0 + chunk]
        Lf = torch.zeros((blk.shape[0], n, n), dtype=torch.float64, device=Lp.device)
        Lf[:, ii, jj] = blk
        Lf[:, jj, ii] = blk
        X = torch.matmul(Lf, C.contiguous())
        Xt = X.transpose(0, 1).reshape(n, -1).contiguous()
        K += Xt @ Xt.t().contiguous()
    return J, K, P


def test_uneven_batch_64_fp64_and_ladder(gpu):
    shapes = uneven_shapes(64, seed=5)
    batch = MoleculeBatch(shapes, device=gpu, seed=3, with_grad_l=False)
    batch.step("fp64", backward=True)
    torch.cuda.synchronize()
    E64 = batch.E.clone()
    dC64 = [d.clone() for d in batch.dC]
    worst_e = worst_c = 0.0
    for i, (n, M, o) in enumerate(shapes):
        C, hl = batch.C[i], batch.h[i][0]
        ii, jj = torch.tril_indices(n, n, device=gpu)
        h = torch.zeros((n, n), dtype=torch.float64, device=gpu)
        h[ii, jj] = hl
        h[jj, ii] = hl
        J, K, P = _torch_jk(batch.L[i], C, n)
        F = h + J - K
        e = torch.sum(P * h).item() + 0.5 * torch.sum(P * (J - K)).item()
        dc = 4.0 * F @ C
        scale = max(1.0, abs(e))
        worst_e = max(worst_e, abs(e - E64[i].item()) / scale)
        worst_c = max(worst_c, torch.max(torch.abs(dc - dC64[i])).item() / max(1.0, torch.max(torch.abs(dc)).item()))
    print(f"64 systems FP64 vs torch: max rel |dE| {worst_e:.3e}, max rel |d dE/dC| {worst_c:.3e}")
    assert worst_e < 1e-11 and worst_c < 1e-10
    plans = batch.tune_ladder(tol=TOL_E)
    batch.step("ladder", backward=True)
    torch.cuda.synchronize()
    dE = torch.max(torch.abs(batch.E - E64)).item()
    dCerr = max(torch.max(torch.abs(a - b)).item() for a, b in zip(batch.dC, dC64))
    lowered = sum(p.j_fp32 < M for p, (_, M, _) in zip(plans, shapes))
    print(f"64 systems ladder: max |dE| {dE:.3e} Eh, max |d dE/dC| {dCerr:.3e}, {lowered} systems use "
          "reduced-precision segments")
    assert dE <= TOL_E
    assert dCerr <= TOL_DC
    batch.close()
