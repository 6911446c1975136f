"""The reader of the reference's Cholesky files (mjk_choinfo_read, host code:
no device needed) against files written in the reference's layout by
tests/_chofiles.py: dimensions, ThrCom, InfVec, and the LT map of every
reduced set rebuilt from CHRED's shell-pair-local indices and ISOSHL, for
shells whose AOs are consecutive and for shells whose AOs are interleaved
(the way p and d components of one centre interleave); error codes for a
missing file, nSym != 1 and an ISOSHL out of range."""
import os

import numpy as np
import pytest

from _chofiles import write_cholesky_files
from openmolcas_amd import MjkError, _lib
from openmolcas_amd.fock import CholeskyFiles

pytestmark = pytest.mark.skipif(not os.path.exists(_lib.LIB_PATH), reason="library not built")


def problem(n, isoshl, nvec, seed, nred=3):
    rng = np.random.default_rng(seed)
    npr = n * (n + 1) // 2
    s0 = np.sort(rng.choice(npr, size=int(0.8 * npr), replace=False))
    sets = [s0]
    for r in range(1, nred):
        sets.append(np.sort(rng.choice(sets[-1], size=int(0.6 * len(sets[-1])), replace=False)))
    vec_set = np.sort(rng.integers(0, nred, nvec))
    vec_lt = []
    for J in range(nvec):
        v = np.zeros(npr)
        v[sets[vec_set[J]]] = rng.standard_normal(len(sets[vec_set[J]]))
        vec_lt.append(v)
    diag = np.zeros(npr)
    diag[s0] = rng.uniform(0.1, 2.0, len(s0))
    return sets, vec_set, vec_lt, diag


@pytest.mark.parametrize("layout,adrvec", [("contiguous", 1), ("interleaved", 2), ("interleaved", 1)])
def test_reader_rebuilds_reduced_sets(tmp_path, layout, adrvec):
    if layout == "contiguous":
        sizes = [1, 3, 1, 5, 3, 1]
        isoshl = np.repeat(np.arange(1, len(sizes) + 1), sizes)
    else:  # s, p(x,y,z) and d components of two shells interleaved
        isoshl = np.array([1, 2, 3, 2, 3, 2, 3, 4, 5, 4, 5, 4, 5, 4, 5, 4, 5, 6])
    n = len(isoshl)
    sets, vec_set, vec_lt, diag = problem(n, isoshl, 17, seed=5)
    maps, vecs, iadr = write_cholesky_files(tmp_path, n, isoshl, sets, vec_set, vec_lt, diag,
                                            adrvec=adrvec, thrcom=3e-7)
    cf = CholeskyFiles(str(tmp_path / "CHORST"), str(tmp_path / "CHRED"), isoshl)
    assert (cf.nbas, cf.numcho, cf.maxred, cf.adrvec) == (n, 17, 3, adrvec)
    assert cf.nshell == int(isoshl.max()) and cf.thrcom == 3e-7
    for r in range(3):
        got = cf.reduced_set(r)
        assert np.array_equal(got, maps[r])           # same elements in file order
        assert np.array_equal(np.sort(got), np.sort(sets[r]))
    assert np.array_equal(cf.infvec(2), vec_set + 1)
    assert np.array_equal(cf.infvec(3), iadr)
    cf.close()


def test_reader_errors(tmp_path):
    isoshl = np.array([1, 2, 2, 2])
    sets, vec_set, vec_lt, diag = problem(4, isoshl, 3, seed=6, nred=1)
    write_cholesky_files(tmp_path, 4, isoshl, sets, vec_set, vec_lt, diag)
    with pytest.raises(MjkError) as e:
        CholeskyFiles(str(tmp_path / "nope"), str(tmp_path / "CHRED"), isoshl)
    assert e.value.rc == 4
    with pytest.raises(MjkError) as e:
        CholeskyFiles(str(tmp_path / "CHORST"), str(tmp_path / "CHRED"), np.array([1, 2, 3, 2]))
    assert e.value.rc == 1                            # shell 3 > nShell
    raw = np.fromfile(tmp_path / "CHORST", dtype="<i8")
    raw[0] = 2                                        # two irreps: not read here
    raw.tofile(tmp_path / "CHORST2")
    with pytest.raises(MjkError) as e:
        CholeskyFiles(str(tmp_path / "CHORST2"), str(tmp_path / "CHRED"), isoshl)
    assert e.value.rc == 1
