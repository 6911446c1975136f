"""Independent numpy evaluation of the LK screening decisions of CHO_LK_SCF /
CHO_LK_CASSCF (nSym = 1), used to pin the oracle's screened exchange.

It is written from the inequalities the reference states, in matrix form
(whole shells and whole batches at once), not from the oracle's loops:

* per batch B of <= 25 vectors of one reduced set, SvShp(ab) =
  max_{J in B} sum_{p in shell pair ab, p in the reduced set} L_pJ^2
  (src/fock_util/cho_getshfull.F90:48-96);
* when screening (first batch, then every nScreen reduced sets):
  Y[:, k] = sqrt|D| |C[:, k]|, MLk(s, k) = max_{l in s} Y(l, k); the kept
  shells of orbital k are those with MLk * max_s MLk >= tau (SCF,
  src/fock_util/cho_lk_scf.F90:602-627) or MLk >= sqrt(tau) (CASSCF,
  src/fock_util/cho_lk_casscf.F90:688);
* X[a, J] for a kept shell of k sums the b-shells with pairs in the reduced
  set and sqrt|SumAClk(b, k) SvShp(ab)| >= thrv (cho_lk_scf.F90:659-702);
  Keep(a) = any such b;
* Faa(a) = max_{i in a} sum_J X[i, J]^2, and the K block (a >= b, both kept
  shells) is added when MLk(a) MLk(b) >= tau, sqrt|Faa(a) Faa(b)| >=
  tau/MaxRed and the Keep flags hold (:749-830);
* D(ab) starts at the integral diagonal and loses sum_J L_abJ^2 after every
  batch (Update), or is the batch's own sum (Estimate).

Returns K (plain LT) per density and the three counters of the oracle.
"""
import numpy as np


def lt_index(n):
    a, b = np.tril_indices(n)
    return a, b


def screened_exchange(n, sets, vred, vecs, csq_list, nocc, fact, tau, thrv, nbassh, diag0,
                      nscreen, maxred, cas_cut, estimate=False, update=True, batch=25):
    """sets: list of LT index arrays; vred: reduced set per vector (ordered);
    vecs: list of vectors on their reduced sets; csq_list[d]: (n, >= nocc[d])
    AO x orbital.  Returns (K (nden, npair) plain LT, stats)."""
    ns = len(nbassh)
    koff = np.concatenate([[0], np.cumsum(nbassh)[:-1]])
    shell = np.repeat(np.arange(ns), nbassh)
    npr = n * (n + 1) // 2
    ia, ib = lt_index(n)
    nden = len(nocc)
    K = np.zeros((nden, npr))
    stats = np.zeros(3, dtype=np.int64)
    diag = np.array(diag0, dtype=np.float64).copy()
    sumac = [np.array([[np.sum(csq_list[d][koff[s]:koff[s] + nbassh[s], k] ** 2) for s in range(ns)]
                       for k in range(nocc[d])]) for d in range(nden)]
    lists = [[None] * nocc[d] for d in range(nden)]
    mlk = [np.zeros((max(nocc[d], 1), ns)) for d in range(nden)]
    doscreen, kscreen = True, 1
    vred = np.asarray(vred)
    for jred in range(vred[0], vred[-1] + 1):
        idx = np.nonzero(vred == jred)[0]
        if len(idx):
            rs = np.asarray(sets[jred])
            shp_of = {}
            for p in rs:
                a, b = ia[p], ib[p]
                key = (max(shell[a], shell[b]), min(shell[a], shell[b]))
                shp_of[key] = shp_of.get(key, 0) + 1
            nb = min(batch, len(idx))
            for b0 in range(0, len(idx), nb):
                bidx = idx[b0:b0 + nb]
                Lrs = np.stack([vecs[j] for j in bidx])          # (jnum, nrs)
                if estimate:
                    diag[:] = 0.0
                    np.add.at(diag, rs, np.sum(Lrs ** 2, axis=0))
                Lsq = np.zeros((len(bidx), n, n))
                Lsq[:, ia[rs], ib[rs]] = Lrs
                Lsq[:, ib[rs], ia[rs]] = Lrs
                # SvShp: max over vectors of the squared norm per shell pair (LT pairs once)
                sv = np.zeros((ns, ns))
                for sa in range(ns):
                    for sb in range(sa + 1):
                        blk = Lsq[:, koff[sa]:koff[sa] + nbassh[sa], koff[sb]:koff[sb] + nbassh[sb]]
                        if sa == sb:
                            tri = np.tril(np.ones((nbassh[sa], nbassh[sa]), dtype=bool))
                            vals = np.sum(blk[:, tri] ** 2, axis=1)
                        else:
                            vals = np.sum(blk ** 2, axis=(1, 2))
                        sv[sa, sb] = sv[sb, sa] = np.max(vals)
                present = np.zeros((ns, ns), dtype=bool)
                for (sa, sb), cnt in shp_of.items():
                    present[sa, sb] = present[sb, sa] = cnt > 0
                if doscreen:
                    dsq = np.zeros((n, n))
                    dsq[ia, ib] = np.sqrt(np.abs(diag))
                    dsq[ib, ia] = np.sqrt(np.abs(diag))
                for d in range(nden):
                    C = csq_list[d][:, :nocc[d]]
                    for k in range(nocc[d]):
                        if doscreen:
                            y = dsq @ np.abs(C[:, k])
                            m = np.array([np.max(y[koff[s]:koff[s] + nbassh[s]]) for s in range(ns)])
                            m = np.maximum(m, 0.0)
                            ok = (m >= np.sqrt(tau[d])) if cas_cut else (m * np.max(m) >= tau[d])
                            lists[d][k] = np.nonzero(ok)[0]
                            mlk[d][k] = m
                            stats[0] += len(lists[d][k])
                        kept = lists[d][k]
                        m = mlk[d][k]
                        X = np.zeros((len(bidx), n))
                        keep = np.zeros(ns, dtype=bool)
                        for sa in kept:
                            use = present[sa] & (np.sqrt(np.abs(sumac[d][k] * sv[sa])) >= thrv[d])
                            use &= np.array(nbassh) > 0
                            keep[sa] = bool(np.any(use))
                            stats[1] += int(np.sum(use))
                            cols = np.isin(shell, np.nonzero(use)[0])
                            rows = slice(koff[sa], koff[sa] + nbassh[sa])
                            X[:, rows] = Lsq[:, rows][:, :, cols] @ C[cols, k]
                        faa = np.zeros(ns)
                        for sa in kept:
                            if keep[sa]:
                                faa[sa] = np.max(np.abs(np.sum(X[:, koff[sa]:koff[sa] + nbassh[sa]] ** 2,
                                                               axis=0)))
                        for sa in kept:
                            for sb in kept:
                                if sa < sb or m[sa] * m[sb] < tau[d]:
                                    continue
                                if not (np.sqrt(np.abs(faa[sa] * faa[sb])) >= tau[d] / maxred):
                                    continue
                                if not (keep[sa] and keep[sb]):
                                    continue
                                stats[2] += 1
                                ra = np.arange(koff[sa], koff[sa] + nbassh[sa])
                                rb = np.arange(koff[sb], koff[sb] + nbassh[sb])
                                blk = X[:, ra].T @ X[:, rb]
                                aa, bb = np.meshgrid(ra, rb, indexing="ij")
                                sel = aa >= bb
                                K[d, aa[sel] * (aa[sel] + 1) // 2 + bb[sel]] += fact[d] * blk[sel]
                doscreen = False
                if update and nscreen > 0:
                    np.subtract.at(diag, rs, np.sum(Lrs ** 2, axis=0))
        doscreen = kscreen == nscreen
        kscreen = 1 if doscreen else kscreen + 1
    return K, stats
