"""Tile screening of the device-engine exchange (mjk_exchange_screened).  The
drop-ins' screening is the reference's LK, tested in test_gpu_lk.py.

* Localized system (vectors and orbitals with Gaussian AO envelopes): the
  screened exchange skips K tiles and half-transform tiles, and its error
  against the unscreened build stays within the bound the skip rules give:
    K tiles:   each skipped (tile, slice) changes K_ab by less than tau
    X tiles:   a zeroed tile of X^J changes K by at most
               3 ||dX^J||_F ||X^J||_F,  ||dX^J||_F <= sqrt(#tiles) thr_x,
               ||X^J||_F <= ||C||_F ||L^J||_F.
* Dense synthetic vectors (no locality): nothing is skipped and the result is
  bit-identical to the unscreened build.
"""
import numpy as np
import pytest
import torch

from openmolcas_amd.fock import JKContext
from openmolcas_amd.synthetic import chol_vectors_np, npair, orbitals_np, vector_scales

pytestmark = pytest.mark.gpu


def localized_system(n, M, o, width, seed):
    rng = np.random.default_rng(seed)
    a = np.arange(n)
    ii, jj = np.tril_indices(n)
    s = vector_scales(M)
    L = np.empty((M, npair(n)))
    for J in range(M):
        g = np.exp(-0.5 * ((a - rng.uniform(0, n)) / width) ** 2)
        A = rng.standard_normal((n, n))
        L[J] = (((A + A.T) / np.sqrt(2.0)) * np.outer(g, g) * s[J])[ii, jj]
    C = np.empty((n, o))
    for k in range(o):
        C[:, k] = np.exp(-0.5 * ((a - (k + 0.5) * n / o) / width) ** 2) * rng.standard_normal(n)
    return L, C


def lfull_norms(L, n):
    """||L^J||_F of the full symmetric matrices."""
    ii, jj = np.tril_indices(n)
    w = np.where(ii == jj, 1.0, 2.0)
    return np.sqrt((L ** 2 * w).sum(axis=1))


def error_bound(stats, tau, thr_x, L, C, n, tile):
    nslices = stats["k_items"]  # every (tile, slice) item could have been skipped
    bound_k = nslices * tau if stats["k_skipped"] else 0.0
    if stats["x_skipped"] == 0:
        return bound_k
    ntile = (n + tile - 1) // tile
    cn = np.linalg.norm(C)
    xn = cn * lfull_norms(L, n)
    bound_x = float(np.sum(3.0 * np.sqrt(ntile) * thr_x * xn))
    return bound_k + bound_x


def test_screening_skips_localized_tiles(gpu):
    n, M, o = 384, 300, 48
    L, C = localized_system(n, M, o, width=10.0, seed=5)
    jk = JKContext(n, M, device=gpu)
    jk.upload_packed(L)
    Ct = torch.from_numpy(C)
    K0 = jk.exchange([Ct]).cpu().numpy()
    tau, thr_x = 1e-9, 1e-9
    K1 = jk.exchange_screened([Ct], tau, thr_x).cpu().numpy()
    st = jk.screening_stats()
    print("localized n=384:", st)
    assert st["k_skipped"] > 0 and st["x_skipped"] > 0
    assert st["k_skipped"] <= st["k_items"] and st["x_skipped"] <= st["x_tiles"]
    err = np.max(np.abs(K1 - K0))
    assert err <= error_bound(st, tau, thr_x, L, C, n, 64), (err, st)
    # thresholds of zero reproduce the unscreened build exactly
    K2 = jk.exchange_screened([Ct], 0.0, 0.0).cpu().numpy()
    assert np.array_equal(K2, K0)
    assert jk.screening_stats()["k_items"] == 0
    jk.close()


def test_screening_dense_vectors_skip_nothing(gpu):
    n, M = 256, 200
    L = chol_vectors_np(n, M, seed=6)
    C = [torch.from_numpy(orbitals_np(n, 40, seed=7)), torch.from_numpy(orbitals_np(n, 9, seed=8))]
    jk = JKContext(n, M, device=gpu)
    jk.upload_packed(L)
    K0 = jk.exchange(C)
    K1 = jk.exchange_screened(C, 1e-10, 1e-10)
    st = jk.screening_stats()
    assert st["k_skipped"] == 0 and st["x_skipped"] == 0 and st["k_items"] > 0
    assert torch.equal(K0, K1)
    K32 = jk.exchange(C, "fp32")
    assert torch.equal(jk.exchange_screened(C, 1e-10, 1e-10, "fp32"), K32)
    jk.close()
