"""CHO_FSCF and CHO_LK_SCF with point-group symmetry (mjk_cho_fscf_sym,
mjk_cho_lk_scf_sym) against the oracle's C1 CHO_FSCF on the same problem
written in the full basis (unscreened LK equals CHO_FSCF with ExFac = FactXI).

Two irreps (nbas 7 + 4): vectors of JSYM = 1 are block diagonal, vectors of
JSYM = 2 couple the irreps (off-diagonal block); each lives in its own reduced
set whose map indexes the C1 lower triangle, as a caller would build from
IndRed/iRS2F.  Orbitals and densities are per irrep.  Tolerance: relative
1e-11 of max|FLT| (FP64 throughout)."""
import numpy as np
import pytest

import oracle
from openmolcas_amd.fock import JKContext
from openmolcas_amd.synthetic import npair

pytestmark = pytest.mark.gpu


def tri(a, b):
    hi, lo = max(a, b), min(a, b)
    return hi * (hi + 1) // 2 + lo


@pytest.mark.parametrize("lk", [False, True])
@pytest.mark.parametrize("nden,exfac", [(1, 0.5), (2, 1.0)])
def test_cho_fscf_symmetry_matches_c1_oracle(gpu, nden, exfac, lk):
    rng = np.random.default_rng(7 + nden)
    nbas = [7, 4]
    off = [0, 7]
    n = sum(nbas)
    npr = npair(n)
    irrep = [0] * 7 + [1] * 4
    # reduced sets: JSYM=1 pairs (same irrep), JSYM=2 pairs (different irreps)
    rs = [[], []]
    for a in range(n):
        for b in range(a + 1):
            rs[0 if irrep[a] == irrep[b] else 1].append(tri(a, b))
    m1, m2 = 9, 6
    V1 = rng.standard_normal((m1, len(rs[0])))
    V2 = rng.standard_normal((m2, len(rs[1])))
    Lfull = np.zeros((m1 + m2, npr))
    Lfull[:m1, rs[0]] = V1
    Lfull[m1:, rs[1]] = V2
    # orbitals per (density, irrep): nforb + niorb
    nforb = [[1, 0], [0, 1]][:nden]
    niorb = [[1, 1], [2, 0]][:nden]
    porb, Cfull = [], []
    for d in range(nden):
        blocks, cols = [], []
        for s in range(2):
            no = nforb[d][s] + niorb[d][s]
            cs = rng.standard_normal((nbas[s], no))
            blocks.append(cs)
            for k in range(no):
                v = np.zeros(n)
                v[off[s]:off[s] + nbas[s]] = cs[:, k]
                cols.append(v)
        porb.append(blocks)
        Cfull.append(np.stack(cols, axis=1))
    # density (alpha + beta or closed shell), folded, per irrep and in C1
    P = sum(c @ c.T for c in Cfull)
    ii, jj = np.tril_indices(n)
    dfull = np.where(ii == jj, 1.0, 2.0) * P[ii, jj]
    ffull0 = np.zeros((nden, npr))
    ntri = sum(b * (b + 1) // 2 for b in nbas)
    dsym = np.zeros(ntri)
    fsym = rng.standard_normal((nden, ntri))
    t = 0
    for s in range(2):
        for a in range(nbas[s]):
            for b in range(a + 1):
                g = tri(off[s] + a, off[s] + b)
                dsym[t] = dfull[g]
                ffull0[:, g] = fsym[:, t]
                t += 1
    st = oracle.CholStore(n, [np.array(rs[0]), np.array(rs[1])],
                          [(0, v) for v in V1] + [(1, v) for v in V2])
    want_full = oracle.cho_fscf(st, ffull0, [c.shape[1] for c in Cfull], Cfull, dfull, exfac)
    # off-block entries of the C1 result stay at zero: the problem is block diagonal
    blk = np.array([irrep[a] == irrep[b] for a in range(n) for b in range(a + 1)])
    assert np.max(np.abs(want_full[:, ~blk])) <= 1e-12 * np.max(np.abs(want_full))
    want = np.zeros_like(fsym)
    t = 0
    for s in range(2):
        for a in range(nbas[s]):
            for b in range(a + 1):
                want[:, t] = want_full[:, tri(off[s] + a, off[s] + b)]
                t += 1
    jk = JKContext(n, m1 + m2, device=gpu)
    jk.set_reduced_set(0, np.array(rs[0], dtype=np.int64))
    jk.set_reduced_set(1, np.array(rs[1], dtype=np.int64))
    jk.upload_vectors(0, 0, V1)
    jk.upload_vectors(1, m1, V2)
    got = fsym.copy()
    if lk:
        jk.cho_lk_scf_sym(nbas, nden, got, nforb, niorb, porb, dsym, exfac, dmpk=0.1, dfmat=1.0)
    else:
        jk.cho_fscf_sym(nbas, nden, got, nforb, niorb, porb, dsym, exfac)
    jk.close()
    assert np.max(np.abs(got - want)) <= 1e-11 * np.max(np.abs(want))
