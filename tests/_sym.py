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


def dense_vectors(p):
    """The problem's vectors as dense symmetric n x n matrices L^J."""
    n = p["n"]
    ii, jj = np.tril_indices(n)
    out = []
    for rs, V in ((p["rs"][0], p["V1"]), (p["rs"][1], p["V2"])):
        for v in V:
            lt = np.zeros(n * (n + 1) // 2)
            lt[np.asarray(rs)] = v
            m = np.zeros((n, n))
            m[ii, jj] = lt
            m[jj, ii] = lt
            out.append(m)
    return np.array(out)


def make_casscf_problem(do_active, seed):
    """CHO_FMCSCF inputs in irrep blocks on the vectors of make_problem, with
    FI/FA from the oracle's C1 CHO_FMCSCF on the embedded problem and (pw|xy)
    formed from the dense vectors and MOs, laid out in blocks (sp, sw,
    sx <= sy) with sp x sw = sx x sy (cho_eval_waxy.F90:86-141): p fastest,
    then w, then the xy pair (LT y(y+1)/2 + x on a diagonal block)."""
    p = make_problem(1, 1.0, seed=seed)
    n = p["n"]
    rng = np.random.default_rng(seed + 1)
    nforb, niorb, naorb, nfro = [1, 0], [1, 1], [2, 1], [1, 0]
    norb = [NBAS[s] - nfro[s] - 1 for s in range(2)]
    cmo = [np.linalg.qr(rng.standard_normal((b, b)))[0] for b in NBAS]
    nfi = [nforb[s] + niorb[s] for s in range(2)]
    Ci = [cmo[s][:, :nfi[s]] for s in range(2)]
    Ca = [cmo[s][:, nfi[s]:nfi[s] + naorb[s]] for s in range(2)]
    occ = [rng.uniform(0.2, 1.8, naorb[s]) for s in range(2)]
    Cm = [Ca[s] * np.sqrt(occ[s]) for s in range(2)]
    nchm = list(naorb) if do_active else [0, 0]
    factxi = -0.5

    def full_cols(blocks):  # irrep column blocks -> n x sum(k) in the full basis
        cols = []
        for s in range(2):
            for k in range(blocks[s].shape[1]):
                v = np.zeros(n)
                v[OFF[s]:OFF[s] + NBAS[s]] = blocks[s][:, k]
                cols.append(v)
        return np.stack(cols, axis=1) if cols else np.zeros((n, 0))

    def fold_lt(m):
        ii, jj = np.tril_indices(n)
        return np.where(ii == jj, 1.0, 2.0) * m[ii, jj]

    CiF, CaF, CmF = full_cols(Ci), full_cols(Ca), full_cols(Cm)
    CoF = full_cols([cmo[s][:, nfro[s]:nfro[s] + norb[s]] for s in range(2)])
    gidx = np.array([tri(OFF[s] + a, OFF[s] + b)
                     for s in range(2) for a in range(NBAS[s]) for b in range(a + 1)])
    dfull = np.stack([fold_lt(2.0 * CiF @ CiF.T), fold_lt(CmF @ CmF.T)])
    fsym0 = rng.standard_normal((2, len(gidx)))
    ffull0 = np.zeros((2, npair(n)))
    ffull0[:, gidx] = fsym0
    st = oracle.CholStore(n, [np.array(p["rs"][0]), np.array(p["rs"][1])],
                          [(0, v) for v in p["V1"]] + [(1, v) for v in p["V2"]])
    cmoF = np.zeros((n, n))
    cmoF[:, :CoF.shape[1]] = CoF
    want_f, _ = oracle.cho_fmcscf(st, ffull0, sum(nfi), sum(naorb), factxi, dfull, do_active,
                                  np.sqrt(2.0) * CiF, sum(nchm), CmF, CaF, cmoF, 0, CoF.shape[1])
    Ld = dense_vectors(p)
    want_w = []
    aoff, ooff = [0, naorb[0]], [0, norb[0]]
    for sp in range(2):
        for sw in range(2):
            for sy in range(2):
                sx = sp ^ sw ^ sy
                if sx > sy:
                    continue
                P = CoF[:, ooff[sp]:ooff[sp] + norb[sp]]
                Wm = CaF[:, aoff[sw]:aoff[sw] + naorb[sw]]
                X = CaF[:, aoff[sx]:aoff[sx] + naorb[sx]]
                Y = CaF[:, aoff[sy]:aoff[sy] + naorb[sy]]
                g = np.einsum("Jab,ap,bw,Jcd,cx,dy->pwxy", Ld, P, Wm, Ld, X, Y)
                for y in range(naorb[sy]):
                    for x in range(y + 1 if sx == sy else naorb[sx]):
                        want_w.append(g[:, :, x, y].reshape(-1, order="F"))
    return {"vectors": p, "nforb": nforb, "niorb": niorb, "naorb": naorb, "nchm": nchm,
            "nfro": nfro, "norb": norb, "factxi": factxi, "cmo": cmo,
            "porb1": [np.sqrt(2.0) * c for c in Ci], "porb2": Cm, "porb3": Ca,
            "dsym": dfull[:, gidx].copy(), "fsym0": fsym0, "want_f": want_f[:, gidx],
            "want_w": np.concatenate(want_w)}
