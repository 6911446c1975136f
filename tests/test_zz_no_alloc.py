"""The allocation-free Coulomb build (mjk_coulomb_noalloc) through the C ABI
against mjk_coulomb (same kernels, so bit-identical J and V), on two calls in
a row, the second with a smaller input, and the rejection of a build whose
scratch exceeds the fixed buffer."""
import ctypes

import numpy as np
import pytest
import torch

from openmolcas_amd import MjkError
from openmolcas_amd.fock import JKContext
from openmolcas_amd.synthetic import chol_vectors_np, fold, orbitals_np

pytestmark = pytest.mark.gpu


def densities(n, nd, seed):
    out = []
    for d in range(nd):
        C = orbitals_np(n, 5 + d, seed + d)
        out.append(fold(C @ C.T))
    return torch.from_numpy(np.stack(out)).to("cuda")


def both(jk, D, want_v):
    nd = D.shape[0]
    j_ref = torch.empty((nd, jk.npair), dtype=torch.float64, device=D.device)
    j_na = torch.full((nd, jk.npair), float("nan"), dtype=torch.float64, device=D.device)
    v_ref = torch.empty((jk.nchol, nd), dtype=torch.float64, device=D.device) if want_v else None
    v_na = torch.full((jk.nchol, nd), float("nan"), dtype=torch.float64, device=D.device) if want_v else None
    p = lambda t: ctypes.c_void_p(t.data_ptr() if t is not None else 0)  # noqa: E731
    jk._check(jk.lib.mjk_coulomb(jk._h, nd, p(D), jk.npair, p(j_ref), jk.npair, p(v_ref)))
    jk._check(jk.lib.mjk_coulomb_noalloc(jk._h, nd, p(D), jk.npair, p(j_na), jk.npair, p(v_na)))
    jk.sync()
    return j_ref, j_na, v_ref, v_na


@pytest.mark.parametrize("want_v", [True, False])
def test_noalloc_coulomb_matches_allocating_entry(gpu, want_v):
    # first call: larger context, two densities
    L1 = chol_vectors_np(300, 500, seed=1)
    jk1 = JKContext(300, 500, device=gpu)
    jk1.upload_packed(L1)
    j_ref, j_na, v_ref, v_na = both(jk1, densities(300, 2, 10), want_v)
    assert torch.equal(j_ref, j_na)
    if want_v:
        assert torch.equal(v_ref, v_na)
    # second call right after: smaller context, one density (the array keeps the
    # first call's contents; nothing in it may leak into this result)
    L2 = chol_vectors_np(90, 130, seed=2)
    jk2 = JKContext(90, 130, device=gpu)
    jk2.upload_packed(L2)
    j_ref, j_na, v_ref, v_na = both(jk2, densities(90, 1, 20), want_v)
    assert torch.equal(j_ref, j_na)
    if want_v:
        assert torch.equal(v_ref, v_na)
    jk1.close()
    jk2.close()


def test_noalloc_coulomb_rejects_oversized_builds(gpu):
    n, M = 1200, 64   # 16 densities: 16 * 720896 * 8 bytes of padded density alone = 92 MiB
    jk = JKContext(n, M, device=gpu)
    jk.upload_packed(np.zeros((M, n * (n + 1) // 2)))
    D = torch.zeros((16, jk.npair), dtype=torch.float64, device=gpu)
    J = torch.zeros((16, jk.npair), dtype=torch.float64, device=gpu)
    with pytest.raises(MjkError) as e:
        jk._check(jk.lib.mjk_coulomb_noalloc(jk._h, 16, ctypes.c_void_p(D.data_ptr()), jk.npair,
                                             ctypes.c_void_p(J.data_ptr()), jk.npair, None))
    assert e.value.rc == 1 and "64 MiB" in str(e.value)
    jk.close()
