"""The FP64 drop-in (mjk_cho_fscf through the C ABI) against golden vectors
computed in exact rational arithmetic (tests/golden/make_golden.py), with the
vectors uploaded per reduced set as the reference's readers deliver them.
Bound: 1e-13 of max|FLT|."""
import numpy as np
import pytest

from _golden import cases, fmcscf_cases, unpack, unpack_fmcscf
from openmolcas_amd.fock import JKContext

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", cases(), ids=lambda c: c["name"])
def test_dropin_matches_exact_golden(gpu, case):
    n, L, Cs, dlt, flt0, exfac, sets, vec_rs, want = unpack(case)
    jk = JKContext(n, L.shape[0], device=gpu)
    for r, s in enumerate(sets):
        jk.set_reduced_set(r, np.array(s, dtype=np.int64))
    for j, r in enumerate(vec_rs):
        jk.upload_vectors(r, j, L[j][np.array(sets[r])].reshape(1, -1))
    got = flt0.copy()
    nd = len(Cs)
    jk.cho_fscf(nd, got, [0] * nd, [C.shape[1] for C in Cs], Cs, dlt, exfac)
    jk.close()
    assert np.max(np.abs(got - want)) <= 1e-13 * np.max(np.abs(want))


@pytest.mark.parametrize("case", fmcscf_cases(), ids=lambda c: c["name"])
def test_fmcscf_dropin_matches_exact_golden(gpu, case):
    n, factxi, nfro, norb, a = unpack_fmcscf(case)
    jk = JKContext(n, a["L_packed"].shape[0], device=gpu)
    jk.upload_packed(a["L_packed"])
    nk, nchm, na = a["porb1"].shape[1], a["porb2"].shape[1], a["porb3"].shape[1]
    flt = a["flt0"].copy()
    w = np.zeros(norb * na * na * (na + 1) // 2)
    jk.cho_fmcscf(flt, 0, nk, na, factxi, a["dlt"], True, a["porb1"], nchm, a["porb2"],
                  a["porb3"], w, a["cmo"], nfro, norb)
    jk.close()
    assert np.max(np.abs(flt - a["flt_expected"])) <= 1e-13 * np.max(np.abs(a["flt_expected"]))
    assert np.max(np.abs(w - a["w_expected"])) <= 1e-13 * np.max(np.abs(a["w_expected"]))
