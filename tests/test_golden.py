"""The oracle against golden vectors computed in exact rational arithmetic
(tests/golden/make_golden.py): CHO_FSCF with one and two densities and with
vectors in two reduced sets.  Bound: 1e-13 of max|FLT| (the inputs are exact
binary fractions, so only the oracle's own FP64 rounding remains)."""
import numpy as np
import pytest

import oracle
from _golden import cases, fmcscf_cases, unpack, unpack_fmcscf


@pytest.mark.parametrize("case", cases(), ids=lambda c: c["name"])
def test_oracle_matches_exact_golden(case):
    n, L, Cs, dlt, flt0, exfac, sets, vec_rs, want = unpack(case)
    vectors = [(r, L[j][np.array(sets[r])]) for j, r in enumerate(vec_rs)]
    st = oracle.CholStore(n, [np.array(s) for s in sets], vectors)
    got = oracle.cho_fscf(st, flt0, [C.shape[1] for C in Cs], Cs, dlt, exfac)
    assert np.max(np.abs(got - want)) <= 1e-13 * np.max(np.abs(want))


@pytest.mark.parametrize("case", fmcscf_cases(), ids=lambda c: c["name"])
def test_oracle_fmcscf_matches_exact_golden(case):
    n, factxi, nfro, norb, a = unpack_fmcscf(case)
    st = oracle.CholStore.full(a["L_packed"])
    nk, nchm, na = a["porb1"].shape[1], a["porb2"].shape[1], a["porb3"].shape[1]
    got, w = oracle.cho_fmcscf(st, a["flt0"], nk, na, factxi, a["dlt"], True, a["porb1"], nchm,
                               a["porb2"], a["porb3"], a["cmo"], nfro, norb)
    assert np.max(np.abs(got - a["flt_expected"])) <= 1e-13 * np.max(np.abs(a["flt_expected"]))
    assert np.max(np.abs(w - a["w_expected"])) <= 1e-13 * np.max(np.abs(a["w_expected"]))
