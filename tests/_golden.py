"""Loader of the exact-arithmetic golden cases (tests/golden/jk_exact.json)."""
import json
import os

import numpy as np

PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "jk_exact.json")


def cases():
    with open(PATH) as fh:
        return json.load(fh)["cases"]


def unpack(c):
    """(n, L packed (M, npair), [C_d], dlt, flt0 (nden, npair), exfac, sets, vec_rs, want)."""
    n = c["nbas"]
    L = np.array(c["L_packed"])
    Cs = [np.array(C) for C in c["C"]]
    npr = n * (n + 1) // 2
    sets = c["reduced_sets"] or [list(range(npr))]
    vec_rs = c["vector_reduced_set"] or [0] * L.shape[0]
    return (n, L, Cs, np.array(c["dlt"]), np.array(c["flt0"]), c["exfac"], sets, vec_rs,
            np.array(c["flt_expected"]))


def fmcscf_cases():
    with open(PATH) as fh:
        return json.load(fh)["fmcscf"]


def unpack_fmcscf(c):
    a = {k: np.array(v) for k, v in c.items()
         if k in ("L_packed", "porb1", "porb2", "porb3", "cmo", "dlt", "flt0", "flt_expected",
                  "w_expected")}
    return c["nbas"], c["factxi"], c["nfro"], c["norb"], a
