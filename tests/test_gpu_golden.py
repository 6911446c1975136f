"""The FP64 drop-in (mjk_cho_fscf through the C ABI) against golden vectors
computed in exact rational arithmetic (tests/golden/make_golden.py), with the
vectors uploaded per reduced set as the reference's readers deliver them.
Bound: 1e-13 of max|FLT|."""
import numpy as np
import pytest

from _golden import cases, unpack
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
