"""Dense numpy brute force for the J/K build -- written from the mathematics,
independent of the oracle's C restatement, so that the two pin each other.

  V_J = sum_ab L^J_ab D_ab,  J = sum_J V_J L^J
  X^J = L^J C,               K = sum_J X^J X^J^T
"""
import numpy as np

from openmolcas_amd.synthetic import pack, unpack, npair


def lfull(L_packed: np.ndarray, n: int) -> np.ndarray:
    M = L_packed.shape[0]
    out = np.zeros((M, n, n))
    i, j = np.tril_indices(n)
    out[:, i, j] = L_packed
    out[:, j, i] = L_packed
    return out


def coulomb(Lf: np.ndarray, D: np.ndarray):
    V = np.einsum("jab,ab->j", Lf, D)
    return np.einsum("j,jab->ab", V, Lf), V


def exchange(Lf: np.ndarray, C: np.ndarray) -> np.ndarray:
    X = np.einsum("jab,bk->jak", Lf, C)
    return np.einsum("jak,jbk->ab", X, X)


def energy_closed_shell(Lf, C, h_sq):
    """E = tr(Ph) + 1/2 tr(P J[P]) - 1/4 tr(P K[P]), P = 2 C C^T."""
    P = 2.0 * C @ C.T
    J, _ = coulomb(Lf, P)
    K = exchange(Lf, np.sqrt(2.0) * C)  # K[P] = sum L P L = sum (L sqrt2 C)(L sqrt2 C)^T
    return np.sum(P * h_sq) + 0.5 * np.sum(P * J) - 0.25 * np.sum(P * K), P, J, K
