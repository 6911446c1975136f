"""Writes the reference's Cholesky files for a synthetic problem (test input).

Layouts, all word-addressed with 8-byte integers and doubles:
  CHORST  src/cholesky_util/cho_wrrstc.F90:30-92
  CHRED   src/cholesky_util/cho_putred.F90:25-44 (cho_putred1.F90): per
          reduced set nnBstRSh(1,1:nnShl), IndRed; set 1 also IndRSh, iSP2F
  CHVEC1  vectors at InfVec(J,3), word (Cho_AdrVec = 1) or 512-byte block
          (Cho_AdrVec = 2) addresses
  CHODIAG src/cholesky_util/cho_iodiag.F90:28-33, reduced-set-1 order
Reduced set 1 is grouped by shell pair (iSP2F ascending) with the
shell-pair-local index LAB of CHO_RSTOF (src/cholesky_util/cho_rstof.F90:
72-80) ascending inside a pair; later sets list positions in set 1.
Returns the LT index of every element of every set in file order and the
vectors in that order, so a reader can be checked against them.
"""
import numpy as np


def itri(i, j):  # 1-based, symmetric
    i, j = max(i, j), min(i, j)
    return i * (i - 1) // 2 + j


def write_cholesky_files(d, n, isoshl, sets_lt, vec_set, vec_lt, diag_lt, adrvec=1, thrcom=1e-6):
    """sets_lt[r]: LT indices of reduced set r (set 0 a superset of all);
    vec_set[J]: 0-based set of vector J (nondecreasing); vec_lt[J]: the
    vector on the full LT; diag_lt: the diagonal on the LT."""
    isoshl = np.asarray(isoshl)
    nshell = int(isoshl.max())
    nbstsh = np.bincount(isoshl, minlength=nshell + 1)[1:]
    ishlso = np.zeros(n, dtype=np.int64)
    cnt = np.zeros(nshell + 1, dtype=np.int64)
    for i in range(n):
        cnt[isoshl[i]] += 1
        ishlso[i] = cnt[isoshl[i]]

    def pair_info(p):
        a = int((np.sqrt(8 * p + 1) - 1) // 2)
        while a * (a + 1) // 2 > p:
            a -= 1
        while (a + 1) * (a + 2) // 2 <= p:
            a += 1
        b = p - a * (a + 1) // 2
        sa, sb = int(isoshl[a]), int(isoshl[b])
        la, lb = int(ishlso[a]), int(ishlso[b])
        if sa == sb:
            lab = itri(la, lb)
        elif sa > sb:
            lab = int(nbstsh[sa - 1]) * (lb - 1) + la
        else:
            lab = int(nbstsh[sb - 1]) * (la - 1) + lb
        return itri(sa, sb), lab

    info1 = {int(p): pair_info(int(p)) for p in sets_lt[0]}
    isp2f = sorted({v[0] for v in info1.values()})
    sp_of = {f: k + 1 for k, f in enumerate(isp2f)}
    nnshl = len(isp2f)
    order1 = sorted(info1, key=lambda p: (info1[p][0], info1[p][1]))
    pos1 = {p: k + 1 for k, p in enumerate(order1)}
    maps, red_words, infred = [], [], []
    adr = 0
    for r, s in enumerate(sets_lt):
        infred.append(adr)
        if r == 0:
            order = order1
        else:
            order = sorted((int(p) for p in s), key=lambda p: (info1[p][0], pos1[p]))
        nn = np.zeros(nnshl, dtype=np.int64)
        for p in order:
            nn[sp_of[info1[p][0]] - 1] += 1
        words = [nn]
        if r == 0:
            words += [np.array([info1[p][1] for p in order], dtype=np.int64),
                      np.array([sp_of[info1[p][0]] for p in order], dtype=np.int64),
                      np.array(isp2f, dtype=np.int64)]
        else:
            words += [np.array([pos1[p] for p in order], dtype=np.int64)]
        blk = np.concatenate(words)
        red_words.append(blk)
        adr += blk.size
        maps.append(np.array(order, dtype=np.int64))
    np.concatenate(red_words).astype("<i8").tofile(f"{d}/CHRED")
    # vectors
    mbl = 8 if adrvec == 1 else 512
    chunks, iadr, at_bytes = [np.zeros(3 * mbl // 8)], [], 3 * mbl  # three units of something else first
    vecs = []
    for J, r in enumerate(vec_set):
        v = np.asarray(vec_lt[J])[maps[r]]
        vecs.append(v)
        iadr.append(at_bytes // mbl)
        nbytes = -(-v.size * 8 // mbl) * mbl
        pad = np.full((nbytes - v.size * 8) // 8, np.nan)
        chunks += [v, pad]
        at_bytes += nbytes
    np.concatenate(chunks).astype("<f8").tofile(f"{d}/CHVEC1")
    nvec = len(vec_set)
    infvec = np.zeros((5, nvec), dtype=np.int64)
    infvec[0] = np.arange(1, nvec + 1)
    infvec[1] = np.asarray(vec_set) + 1
    infvec[2] = iadr
    head = np.array([1, nshell, nnshl, 0, n, 0, adrvec], dtype=np.int64)
    dscr = np.array([thrcom, 1e-9, 1.0, 1.0, 1e-2, -1e-40, -1e-8, -1e-6])
    rest = np.concatenate([[len(sets_lt)], infred, [nvec], infvec.reshape(-1)]).astype(np.int64)
    with open(f"{d}/CHORST", "wb") as f:
        head.astype("<i8").tofile(f)
        dscr.astype("<f8").tofile(f)
        rest.astype("<i8").tofile(f)
    np.asarray(diag_lt)[maps[0]].astype("<f8").tofile(f"{d}/CHODIAG")
    return maps, vecs, np.array(iadr, dtype=np.int64)


def write_cholesky_files_sym(d, nbas_sym, isoshl, sets_lt, vec_sym, vec_set, vec_lt, diag_lt, adrvec=1,
                             thrcom=1e-6):
    """The same files for nSym = len(nbas_sym) irreps, SOs numbered irrep by
    irrep: sets_lt[r] are C1 LT indices of reduced set r over all irreps;
    vec_sym[J] (0-based irrep, the product of the irreps of every pair the
    vector touches) and vec_set[J] (0-based set) per vector, irreps
    nondecreasing and sets nondecreasing inside an irrep.  CHRED lists each
    set irrep by irrep, then by shell pair (cho_setredind.F90), with
    nnBstRSh(1:nSym, 1:nnShl) irrep-fastest; CHORST carries nBas(1:nSym) and
    per irrep NumCho and InfVec (cho_wrrstc.F90:62-74); irrep s's vectors go
    to CHVEC<s+1>.  Returns ({(irrep, set): LT indices in file order}, the
    vectors per irrep in file order, the InfVec(:,3) addresses per irrep)."""
    isoshl = np.asarray(isoshl)
    nsym = len(nbas_sym)
    n = int(sum(nbas_sym))
    so_sym = np.repeat(np.arange(nsym), nbas_sym)
    nshell = int(isoshl.max())
    nbstsh = np.bincount(isoshl, minlength=nshell + 1)[1:]
    ishlso = np.zeros(n, dtype=np.int64)
    cnt = np.zeros(nshell + 1, dtype=np.int64)
    for i in range(n):
        cnt[isoshl[i]] += 1
        ishlso[i] = cnt[isoshl[i]]

    def pair_info(p):
        a = int((np.sqrt(8 * p + 1) - 1) // 2)
        while a * (a + 1) // 2 > p:
            a -= 1
        while (a + 1) * (a + 2) // 2 <= p:
            a += 1
        b = p - a * (a + 1) // 2
        sa, sb = int(isoshl[a]), int(isoshl[b])
        la, lb = int(ishlso[a]), int(ishlso[b])
        if sa == sb:
            lab = itri(la, lb)
        elif sa > sb:
            lab = int(nbstsh[sa - 1]) * (lb - 1) + la
        else:
            lab = int(nbstsh[sb - 1]) * (la - 1) + lb
        return int(so_sym[a] ^ so_sym[b]), itri(sa, sb), lab

    info1 = {int(p): pair_info(int(p)) for p in sets_lt[0]}
    isp2f = sorted({v[1] for v in info1.values()})
    sp_of = {f: k + 1 for k, f in enumerate(isp2f)}
    nnshl = len(isp2f)
    order1 = sorted(info1, key=lambda p: info1[p])
    pos1 = {p: k + 1 for k, p in enumerate(order1)}
    maps, red_words, infred = {}, [], []
    adr = 0
    for r, s in enumerate(sets_lt):
        infred.append(adr)
        if r == 0:
            order = order1
        else:
            order = sorted((int(p) for p in s), key=lambda p: (info1[p][0], info1[p][1], pos1[p]))
        nn = np.zeros((nnshl, nsym), dtype=np.int64)  # row-major (sp, irrep): irrep fastest on disk
        for p in order:
            nn[sp_of[info1[p][1]] - 1, info1[p][0]] += 1
        words = [nn.reshape(-1)]
        if r == 0:
            words += [np.array([info1[p][2] for p in order], dtype=np.int64),
                      np.array([sp_of[info1[p][1]] for p in order], dtype=np.int64),
                      np.array(isp2f, dtype=np.int64)]
        else:
            words += [np.array([pos1[p] for p in order], dtype=np.int64)]
        blk = np.concatenate(words)
        red_words.append(blk)
        adr += blk.size
        for sym in range(nsym):
            maps[(sym, r)] = np.array([p for p in order if info1[p][0] == sym], dtype=np.int64)
    np.concatenate(red_words).astype("<i8").tofile(f"{d}/CHRED")
    mbl = 8 if adrvec == 1 else 512
    vecs, iadrs, infvecs = [], [], []
    for sym in range(nsym):
        js = [J for J in range(len(vec_sym)) if vec_sym[J] == sym]
        chunks, iadr, at_bytes, vs = [np.zeros(2 * mbl // 8)], [], 2 * mbl, []
        for J in js:
            v = np.asarray(vec_lt[J])[maps[(sym, vec_set[J])]]
            vs.append(v)
            iadr.append(at_bytes // mbl)
            nbytes = -(-max(v.size, 1) * 8 // mbl) * mbl
            chunks += [v, np.full((nbytes - v.size * 8) // 8, np.nan)]
            at_bytes += nbytes
        np.concatenate(chunks).astype("<f8").tofile(f"{d}/CHVEC{sym + 1}")
        iv = np.zeros((5, len(js)), dtype=np.int64)
        iv[0] = np.arange(1, len(js) + 1)
        iv[1] = [vec_set[J] + 1 for J in js]
        iv[2] = iadr
        vecs.append(vs)
        iadrs.append(np.array(iadr, dtype=np.int64))
        infvecs.append(np.concatenate([[len(js)], iv.reshape(-1) if js else []]).astype(np.int64))
    head = np.concatenate([[nsym, nshell, nnshl, 0], nbas_sym, [0, adrvec]]).astype(np.int64)
    dscr = np.array([thrcom, 1e-9, 1.0, 1.0, 1e-2, -1e-40, -1e-8, -1e-6])
    rest = np.concatenate([[len(sets_lt)], infred] + infvecs).astype(np.int64)
    with open(f"{d}/CHORST", "wb") as f:
        head.astype("<i8").tofile(f)
        dscr.astype("<f8").tofile(f)
        rest.astype("<i8").tofile(f)
    np.asarray(diag_lt)[order1].astype("<f8").tofile(f"{d}/CHODIAG")
    return maps, vecs, iadrs
