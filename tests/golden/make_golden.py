"""Writes jk_exact.json: small CHO_FSCF cases whose expected Fock matrices are
computed in exact rational arithmetic (fractions.Fraction) from the closed
forms, independently of the oracle and of floating point:

    V_J  = sum_p L^J_p DLT_p                      (DLT folded: off-diagonal doubled)
    J_ab = sum_J L^J_ab V_J
    K_ab = sum_J sum_k (L^J C)_ak (L^J C)_bk       (L^J the full symmetric matrix)
    FLT  = FLT0 + J - ExFac * K                   (src/fock_util/cho_fscf.F90:64-65)

All inputs are small integers or halves, so every expected value is an exact
rational, stored as a float (and as its exact fraction for reference).
Cases: one density; two densities sharing DLT (UHF-like, nDen = 2); and a
store with two reduced sets (the second covering half of the pairs), where the
missing pairs of a vector are zero.  Plus one CHO_FMCSCF case (inactive and
active Fock matrices and the (pw|xy) block, see fmcscf_case).

Run:  python tests/golden/make_golden.py
"""
import json
import os
from fractions import Fraction as Fr

HERE = os.path.dirname(os.path.abspath(__file__))


def npair(n):
    return n * (n + 1) // 2


def lt_index(a, b):
    hi, lo = max(a, b), min(a, b)
    return hi * (hi + 1) // 2 + lo


def full(vec, n):
    return [[vec[lt_index(a, b)] for b in range(n)] for a in range(n)]


def fscf(n, Ls, Cs, dlt, flt0, exfac):
    npr = npair(n)
    V = [sum(L[p] * dlt[p] for p in range(npr)) for L in Ls]
    J = [sum(Ls[j][p] * V[j] for j in range(len(Ls))) for p in range(npr)]
    out = []
    for d, C in enumerate(Cs):
        o = len(C[0]) if C else 0
        K = [Fr(0)] * npr
        for L in Ls:
            Lf = full(L, n)
            X = [[sum(Lf[a][c] * C[c][k] for c in range(n)) for k in range(o)] for a in range(n)]
            for a in range(n):
                for b in range(a + 1):
                    K[lt_index(a, b)] += sum(X[a][k] * X[b][k] for k in range(o))
        out.append([flt0[d][p] + J[p] - exfac * K[p] for p in range(npr)])
    return out


def fold(P, n):
    return [P[a][b] * (1 if a == b else 2) for a in range(n) for b in range(a + 1)]


def case(name, n, Ls, Cs, flt0, exfac, rsets=None, vec_rs=None):
    P = [[sum(C[a][k] * C[b][k] for C in Cs[:1] for k in range(len(C[0]))) for b in range(n)]
         for a in range(n)]
    dlt = fold(P, n)
    want = fscf(n, Ls, Cs, dlt, flt0, exfac)
    f = lambda v: [float(x) for x in v]
    return {
        "name": name, "nbas": n, "nchol": len(Ls), "exfac": float(exfac),
        "L_packed": [f(L) for L in Ls],
        "C": [[f(r) for r in C] for C in Cs],
        "dlt": f(dlt), "flt0": [f(x) for x in flt0],
        "reduced_sets": rsets, "vector_reduced_set": vec_rs,
        "flt_expected": [f(x) for x in want],
        "flt_expected_exact": [[str(x) for x in w] for w in want],
    }


def main():
    n = 4
    npr = npair(n)
    L1 = [Fr(v) for v in (2, -1, 3, 0, 1, -2, 1, 0, 2, -1)]
    L2 = [Fr(v) for v in (1, 0, -1, 2, 1, 0, -1, 1, 0, 3)]
    L3 = [Fr(v, 2) for v in (1, 1, 0, -1, 2, 1, 0, 3, -2, 1)]
    C1 = [[Fr(1), Fr(0)], [Fr(1), Fr(1)], [Fr(0), Fr(-1)], [Fr(2), Fr(1)]]
    C2 = [[Fr(0)], [Fr(1)], [Fr(-1)], [Fr(1)]]
    flt0 = [[Fr(p % 3) - 1 for p in range(npr)], [Fr(1, 2)] * npr]
    cases = [
        case("one density", n, [L1, L2, L3], [C1], flt0[:1], Fr(1, 2)),
        case("two densities", n, [L1, L2, L3], [C1, C2], flt0, Fr(1)),
    ]
    # reduced sets: set 0 = all pairs; set 1 = pairs 0, 2, 4, 6, 8 (vector 2 lives there)
    rs1 = [0, 2, 4, 6, 8]
    L3r = [L3[p] if p in rs1 else Fr(0) for p in range(npr)]
    c = case("two reduced sets", n, [L1, L2, L3r], [C1], flt0[:1], Fr(1, 2),
             rsets=[list(range(npr)), rs1], vec_rs=[0, 0, 1])
    cases.append(c)
    with open(os.path.join(HERE, "jk_exact.json"), "w") as fh:
        json.dump({"source": "tests/golden/make_golden.py (exact rational arithmetic)", "cases": cases,
                   "fmcscf": [fmcscf_case(n, [L1, L2, L3])]}, fh, indent=1)


def kmat(n, Ls, C):
    """K[C] = sum_J L^J C C^T L^J, packed."""
    o = len(C[0])
    K = [Fr(0)] * npair(n)
    for L in Ls:
        Lf = full(L, n)
        X = [[sum(Lf[a][c] * C[c][k] for c in range(n)) for k in range(o)] for a in range(n)]
        for a in range(n):
            for b in range(a + 1):
                K[lt_index(a, b)] += sum(X[a][k] * X[b][k] for k in range(o))
    return K


def jmat(n, Ls, dlt):
    V = [sum(L[p] * dlt[p] for p in range(npair(n))) for L in Ls]
    return [sum(Ls[j][p] * V[j] for j in range(len(Ls))) for p in range(npair(n))]


def fmcscf_case(n, Ls):
    """CHO_FMCSCF (src/fock_util/cho_fmcscf.F90): FI = FLT0(1) + J[DI] + FactXI K[POrb(1)],
    FA = FLT0(2) + J[DA] - 1/2 K[POrb(2)], and W_PWXY[(xy, w, p)] = sum_J (pw|J)(xy|J) with
    (pw|J) = sum_ab L^J_ab CMO(a, nfro+p) Ash(b, w), (xy|J) = sum_ab L^J_ab Ash(a, x) Ash(b, y),
    xy = y(y+1)/2 + x (x <= y), stored xy-major, then w, then p."""
    npr = npair(n)
    Ci = [[Fr(1)], [Fr(-1)], [Fr(0)], [Fr(2)]]
    Cm = [[Fr(1), Fr(0)], [Fr(0), Fr(1, 2)], [Fr(1), Fr(1)], [Fr(0), Fr(-1)]]
    Ca = [[Fr(0), Fr(1)], [Fr(1), Fr(0)], [Fr(1), Fr(-1)], [Fr(0), Fr(1)]]
    cmo = [[Fr(v) for v in row] for row in ((1, 0, 1, 0), (0, 1, 1, -1), (1, 1, 0, 2), (0, -1, 1, 1))]
    nfro, norb, na = 1, 3, 2
    factxi = Fr(-1, 2)
    DI = fold([[2 * sum(Ci[a][k] * Ci[b][k] for k in range(1)) for b in range(n)] for a in range(n)], n)
    DA = fold([[sum(Cm[a][k] * Cm[b][k] for k in range(2)) for b in range(n)] for a in range(n)], n)
    flt0 = [[Fr(p, 4) for p in range(npr)], [Fr(-1)] * npr]
    KI, KA = kmat(n, Ls, Ci), kmat(n, Ls, Cm)
    JI, JA = jmat(n, Ls, DI), jmat(n, Ls, DA)
    FI = [flt0[0][p] + JI[p] + factxi * KI[p] for p in range(npr)]
    FA = [flt0[1][p] + JA[p] - Fr(1, 2) * KA[p] for p in range(npr)]
    W = []
    Lfs = [full(L, n) for L in Ls]
    for y in range(na):
        for x in range(y + 1):
            xyJ = [sum(Lf[a][b] * Ca[a][x] * Ca[b][y] for a in range(n) for b in range(n)) for Lf in Lfs]
            for w in range(na):
                for pp in range(norb):
                    W.append(sum(xyJ[j] * sum(Lfs[j][a][b] * cmo[a][nfro + pp] * Ca[b][w]
                                               for a in range(n) for b in range(n))
                                 for j in range(len(Ls))))
    f = lambda v: [float(x) for x in v]
    return {
        "name": "fmcscf inactive + active, (pw|xy)", "nbas": n, "nchol": len(Ls),
        "L_packed": [f(L) for L in Ls], "factxi": float(factxi), "nfro": nfro, "norb": norb,
        "porb1": [f(r) for r in Ci], "porb2": [f(r) for r in Cm], "porb3": [f(r) for r in Ca],
        "cmo": [f(r) for r in cmo], "dlt": [f(DI), f(DA)], "flt0": [f(x) for x in flt0],
        "flt_expected": [f(FI), f(FA)], "w_expected": f(W),
        "flt_expected_exact": [[str(x) for x in FI], [str(x) for x in FA]],
        "w_expected_exact": [str(x) for x in W],
    }


if __name__ == "__main__":
    main()
