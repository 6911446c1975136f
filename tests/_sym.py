"""A two-irrep CHO_FSCF problem (nbas 7 + 4) for the symmetry drop-ins and its
C1 reference from the oracle.  Vectors of JSYM = 1 are block diagonal, vectors
of JSYM = 2 couple the irreps; each symmetry has its own reduced set whose map
indexes the C1 lower triangle."""
import numpy as np

import oracle
from openmolcas_amd.synthetic import npair

NBAS = [7, 4]
OFF = [0, 7]


def tri(a, b):
    hi, lo = max(a, b), min(a, b)
    return hi * (hi + 1) // 2 + lo


def make_problem(nden, exfac, seed):
    rng = np.random.default_rng(seed)
    n = sum(NBAS)
    npr = npair(n)
    irrep = [0] * NBAS[0] + [1] * NBAS[1]
    rs = [[], []]
    for a in range(n):
        for b in range(a + 1):
            rs[0 if irrep[a] == irrep[b] else 1].append(tri(a, b))
    m1, m2 = 9, 6
    V1 = rng.standard_normal((m1, len(rs[0])))
    V2 = rng.standard_normal((m2, len(rs[1])))
    nforb = [[1, 0], [0, 1]][:nden]
    niorb = [[1, 1], [2, 0]][:nden]
    porb, Cfull = [], []
    for d in range(nden):
        blocks, cols = [], []
        for s in range(2):
            no = nforb[d][s] + niorb[d][s]
            cs = rng.standard_normal((NBAS[s], no))
            blocks.append(cs)
            for k in range(no):
                v = np.zeros(n)
                v[OFF[s]:OFF[s] + NBAS[s]] = cs[:, k]
                cols.append(v)
        porb.append(blocks)
        Cfull.append(np.stack(cols, axis=1))
    P = sum(c @ c.T for c in Cfull)
    ii, jj = np.tril_indices(n)
    dfull = np.where(ii == jj, 1.0, 2.0) * P[ii, jj]
    ntri = sum(b * (b + 1) // 2 for b in NBAS)
    dsym = np.zeros(ntri)
    fsym = rng.standard_normal((nden, ntri))
    ffull0 = np.zeros((nden, npr))
    gidx = []
    for s in range(2):
        for a in range(NBAS[s]):
            for b in range(a + 1):
                gidx.append(tri(OFF[s] + a, OFF[s] + b))
    gidx = np.array(gidx)
    dsym[:] = dfull[gidx]
    ffull0[:, gidx] = fsym
    st = oracle.CholStore(n, [np.array(rs[0]), np.array(rs[1])],
                          [(0, v) for v in V1] + [(1, v) for v in V2])
    want_full = oracle.cho_fscf(st, ffull0, [c.shape[1] for c in Cfull], Cfull, dfull, exfac)
    blk = np.array([irrep[a] == irrep[b] for a in range(n) for b in range(a + 1)])
    return {"n": n, "rs": rs, "V1": V1, "V2": V2, "nforb": nforb, "niorb": niorb, "porb": porb,
            "dsym": dsym, "fsym": fsym, "want_full": want_full, "offblock": ~blk,
            "want": want_full[:, gidx]}
