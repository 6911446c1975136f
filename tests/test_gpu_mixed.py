"""Mixed-precision exchange (FP32 MFMA, FP64 reduction over J) against the
FP64 oracle.  Tolerance: the FP32 path rounds L and C to FP32 (relative 6e-8)
and accumulates at most 512 products in FP32 before folding into FP64, so K
agrees with the FP64 result to a relative max-abs error of 2e-6."""
import numpy as np
import pytest
import torch

import oracle
from openmolcas_amd.fock import JKContext
from openmolcas_amd.synthetic import chol_vectors_np, orbitals_np

pytestmark = pytest.mark.gpu

FP32_RTOL = 2e-6
# split-BF16: inputs carried to 16 significant bits (hi + lo), lo*lo dropped,
# FP32 accumulation: relative max-abs error 1e-4 of max|K|
B3_RTOL = 1e-4


def oracle_K(L, C):
    st = oracle.CholStore.full(L)
    return oracle.cho_fscf(st, np.zeros((1, L.shape[1])), [C.shape[1]], [C],
                           np.zeros(L.shape[1]), -1.0)[0]


@pytest.mark.parametrize("n,M,occs", [(64, 33, [16, 1]), (200, 40, [64, 13]), (300, 64, [150]),
                                      (160, 30, [120, 80]), (150, 21, [70]), (170, 17, [100])])
def test_exchange_fp32_matches_fp64_oracle(gpu, n, M, occs):
    L = chol_vectors_np(n, M, seed=3 * n + M)
    Cs = [orbitals_np(n, o, seed=500 + i) for i, o in enumerate(occs)]
    ctx = JKContext(n, M, device=gpu)
    ctx.upload_packed(L)
    K32 = ctx.exchange([torch.from_numpy(c) for c in Cs], precision="fp32").cpu().numpy()
    K64 = ctx.exchange([torch.from_numpy(c) for c in Cs], precision="fp64").cpu().numpy()
    for d, C in enumerate(Cs):
        ref = oracle_K(L, C)
        scale = np.max(np.abs(ref))
        assert np.max(np.abs(K64[d] - ref)) / scale < 1e-12
        err32 = np.max(np.abs(K32[d] - ref)) / scale
        assert err32 < FP32_RTOL, err32
        assert err32 > 0.0  # the FP32 path really ran in FP32


def test_fp32_shadow_refreshed_after_upload(gpu):
    n, M = 48, 20
    L = chol_vectors_np(n, M, seed=1)
    C = orbitals_np(n, 8, seed=2)
    ctx = JKContext(n, M, device=gpu)
    ctx.upload_packed(L)
    k1 = ctx.exchange([torch.from_numpy(C)], precision="fp32").cpu().numpy()
    ctx.upload_packed(2.0 * L)  # the FP32 copy of L must follow the new vectors
    k2 = ctx.exchange([torch.from_numpy(C)], precision="fp32").cpu().numpy()
    assert np.allclose(k2, 4.0 * k1, rtol=1e-6, atol=1e-9 * np.max(np.abs(k1)))


@pytest.mark.parametrize("n,M,occs", [(64, 33, [16, 1]), (200, 40, [64, 13]), (300, 64, [150]),
                                      (160, 30, [120, 80]), (150, 21, [70]), (170, 17, [100])])
def test_exchange_bf16x3_matches_fp64_oracle(gpu, n, M, occs):
    L = chol_vectors_np(n, M, seed=5 * n + M)
    Cs = [orbitals_np(n, o, seed=700 + i) for i, o in enumerate(occs)]
    ctx = JKContext(n, M, device=gpu)
    ctx.upload_packed(L)
    K3 = ctx.exchange([torch.from_numpy(c) for c in Cs], precision="bf16x3").cpu().numpy()
    for d, C in enumerate(Cs):
        ref = oracle_K(L, C)
        err = np.max(np.abs(K3[d] - ref)) / np.max(np.abs(ref))
        assert err < B3_RTOL, err
