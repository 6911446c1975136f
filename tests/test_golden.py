"""The oracle against golden vectors computed in exact rational arithmetic
(tests/golden/make_golden.py): CHO_FSCF with one and two densities and with
vectors in two reduced sets.  Bound: 1e-13 of max|FLT| (the inputs are exact
binary fractions, so only the oracle's own FP64 rounding remains)."""
import numpy as np
import pytest

import oracle
from _golden import cases, unpack


@pytest.mark.parametrize("case", cases(), ids=lambda c: c["name"])
def test_oracle_matches_exact_golden(case):
    n, L, Cs, dlt, flt0, exfac, sets, vec_rs, want = unpack(case)
    vectors = [(r, L[j][np.array(sets[r])]) for j, r in enumerate(vec_rs)]
    st = oracle.CholStore(n, [np.array(s) for s in sets], vectors)
    got = oracle.cho_fscf(st, flt0, [C.shape[1] for C in Cs], Cs, dlt, exfac)
    assert np.max(np.abs(got - want)) <= 1e-13 * np.max(np.abs(want))
