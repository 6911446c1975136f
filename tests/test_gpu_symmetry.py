"""CHO_FSCF and CHO_LK_SCF with point-group symmetry (mjk_cho_fscf_sym,
mjk_cho_lk_scf_sym) against the oracle's C1 CHO_FSCF on the same problem
written in the full basis (unscreened LK equals CHO_FSCF with ExFac = FactXI).
Two irreps (nbas 7 + 4), vectors of JSYM = 1 and 2 in their own reduced sets
(tests/_sym.py).  Tolerance: relative 1e-11 of max|FLT| (FP64 throughout)."""
import numpy as np
import pytest

from _sym import NBAS, make_problem
from openmolcas_amd.fock import JKContext

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("lk", [False, True])
@pytest.mark.parametrize("nden,exfac", [(1, 0.5), (2, 1.0)])
def test_cho_fscf_symmetry_matches_c1_oracle(gpu, nden, exfac, lk):
    p = make_problem(nden, exfac, seed=7 + nden)
    # the problem is block diagonal: the C1 result has no off-irrep entries
    assert np.max(np.abs(p["want_full"][:, p["offblock"]])) <= 1e-12 * np.max(np.abs(p["want_full"]))
    m1 = p["V1"].shape[0]
    jk = JKContext(p["n"], m1 + p["V2"].shape[0], device=gpu)
    jk.set_reduced_set(0, np.array(p["rs"][0], dtype=np.int64))
    jk.set_reduced_set(1, np.array(p["rs"][1], dtype=np.int64))
    jk.upload_vectors(0, 0, p["V1"])
    jk.upload_vectors(1, m1, p["V2"])
    got = p["fsym"].copy()
    if lk:
        jk.cho_lk_scf_sym(NBAS, nden, got, p["nforb"], p["niorb"], p["porb"], p["dsym"], exfac,
                          dmpk=0.1, dfmat=1.0)
    else:
        jk.cho_fscf_sym(NBAS, nden, got, p["nforb"], p["niorb"], p["porb"], p["dsym"], exfac)
    jk.close()
    want = p["want"]
    assert np.max(np.abs(got - want)) <= 1e-11 * np.max(np.abs(want))
