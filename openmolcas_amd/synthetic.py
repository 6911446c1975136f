"""Seeded synthetic inputs for the J/K build.

* Cholesky vectors: L^J symmetric n x n with i.i.d. N(0,1) entries symmetrised
  as (A + A^T)/sqrt(2) -- off-diagonal N(0,1), diagonal N(0,2) -- scaled by
  s_J = 10^(-6 J/(M-1)) so that vector norms decay geometrically from 1 to
  1e-6, stored packed (LT, a >= b, index a(a+1)/2 + b).
* Orbitals: QR of an N(0,1) n x o matrix (orthonormal columns).
* One-electron matrix: h = sym N(0,1)/n, packed LT.

numpy generators (PCG64) serve the small parity cases; the torch generators
produce the same distributions directly in HBM for benchmark sizes.
"""
from __future__ import annotations

import math

import numpy as np


def npair(n: int) -> int:
    return n * (n + 1) // 2


def diag_index(n: int) -> np.ndarray:
    a = np.arange(n)
    return a * (a + 1) // 2 + a


def vector_scales(M: int) -> np.ndarray:
    if M == 1:
        return np.ones(1)
    return 10.0 ** (-6.0 * np.arange(M) / (M - 1))


def chol_vectors_np(n: int, M: int, seed: int) -> np.ndarray:
    """(M, n(n+1)/2) float64 packed vectors."""
    rng = np.random.Generator(np.random.PCG64(seed))
    L = rng.standard_normal((M, npair(n)))
    L[:, diag_index(n)] *= math.sqrt(2.0)
    L *= vector_scales(M)[:, None]
    return L


def orbitals_np(n: int, o: int, seed: int) -> np.ndarray:
    """(n, o) with orthonormal columns."""
    rng = np.random.Generator(np.random.PCG64(seed))
    q, _ = np.linalg.qr(rng.standard_normal((n, o)))
    return np.ascontiguousarray(q)


def hcore_np(n: int, seed: int) -> np.ndarray:
    rng = np.random.Generator(np.random.PCG64(seed))
    a = rng.standard_normal((n, n))
    s = (a + a.T) / (2.0 * n)
    return pack(s)


def pack(a: np.ndarray) -> np.ndarray:
    """Square symmetric -> LT packed (plain, not folded)."""
    n = a.shape[0]
    i, j = np.tril_indices(n)
    # row-major lower triangle is exactly a(a+1)/2 + b ordering
    return np.ascontiguousarray(a[i, j])


def unpack(p: np.ndarray, n: int) -> np.ndarray:
    """LT packed (plain) -> square symmetric."""
    out = np.zeros((n, n))
    i, j = np.tril_indices(n)
    out[i, j] = p
    out[j, i] = p
    return out


def fold(a: np.ndarray) -> np.ndarray:
    """Square symmetric density -> folded LT (off-diagonal doubled)."""
    p = pack(a) * 2.0
    p[diag_index(a.shape[0])] *= 0.5
    return p


# ------------------------------------------------------------------ torch (HBM)
def chol_vectors_torch(n: int, M: int, seed: int, device="cuda", dtype=None, chunk: int = 256,
                       out=None):
    """(M, n(n+1)/2) packed vectors generated on the device in chunks (into
    `out` when given, e.g. a context's resident store)."""
    import torch

    dtype = dtype or torch.float64
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    npr = npair(n)
    L = torch.empty((M, npr), dtype=dtype, device=device) if out is None else out
    scales = torch.tensor(vector_scales(M), dtype=dtype, device=device)
    di = torch.tensor(diag_index(n), dtype=torch.long, device=device)
    for j0 in range(0, M, chunk):
        j1 = min(M, j0 + chunk)
        blk = torch.randn((j1 - j0, npr), generator=g, dtype=dtype, device=device)
        blk[:, di] *= math.sqrt(2.0)
        blk *= scales[j0:j1, None]
        L[j0:j1] = blk
    return L


def orbitals_torch(n: int, o: int, seed: int, device="cuda"):
    import torch

    g = torch.Generator(device=device)
    g.manual_seed(seed)
    a = torch.randn((n, o), generator=g, dtype=torch.float64, device=device)
    q, _ = torch.linalg.qr(a)
    return q.contiguous()
