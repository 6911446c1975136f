"""Shell-structured, spatially local inputs for the LK screening tests.

A chain of "atoms" on a line, each carrying shells of 1, 3 or 5 functions
(s, p, d); an AO sits at its atom's position.  The Cholesky vectors are
local the way real ones are: vector J is centred at c_J and
    L^J_ab = exp(-alpha (x_a - x_b)^2) exp(-gamma ((x_a + x_b)/2 - c_J)^2) g_abJ s_J
with g ~ N(0,1) and s_J decaying over J, so pairs of distant AOs and pairs
far from c_J are small.  The integral diagonal is D_ab = sum_J (L^J_ab)^2
(the exact diagonal of a full decomposition), and the orbitals are localized
around random points.  The vectors can be split over reduced sets whose
pair lists shrink to the pairs of largest diagonal, as the reference's
decomposition produces them.
"""
import numpy as np


def molecule(natom=6, seed=0, spacing=1.6, alpha=0.35, gamma=0.25, nvec=70, nred=3,
             shells_per_atom=(1, 3, 1), decay=4.0):
    rng = np.random.default_rng(seed)
    nbassh, pos = [], []
    for a in range(natom):
        for sz in shells_per_atom:
            nbassh.append(int(sz))
            pos += [a * spacing] * int(sz)
    pos = np.array(pos)
    n = len(pos)
    ia, ib = np.tril_indices(n)
    npr = len(ia)
    xa, xb = pos[ia], pos[ib]
    over = np.exp(-alpha * (xa - xb) ** 2)
    mid = 0.5 * (xa + xb)
    centres = rng.uniform(pos.min(), pos.max(), nvec)
    scale = 10.0 ** (-decay * np.arange(nvec) / max(nvec - 1, 1))
    L = np.empty((nvec, npr))
    for j in range(nvec):
        L[j] = over * np.exp(-gamma * (mid - centres[j]) ** 2) * rng.standard_normal(npr) * scale[j]
    diag = np.sum(L * L, axis=0)
    # reduced sets: all pairs, then pairs of ever larger diagonal
    order = np.argsort(-diag)
    sets, vred = [], []
    per = -(-nvec // nred)
    for r in range(nred):
        keep = np.sort(order[: max(1, npr * (nred - r) // nred)])
        sets.append(keep.astype(np.int64))
        vred += [r] * len(range(r * per, min(nvec, (r + 1) * per)))
    vecs = [L[j, sets[vred[j]]] for j in range(nvec)]
    Lfull = np.zeros_like(L)  # the vectors as stored on their reduced sets, back on the LT
    for j in range(nvec):
        Lfull[j, sets[vred[j]]] = vecs[j]
    return dict(n=n, nbassh=nbassh, pos=pos, sets=sets, vred=vred, vecs=vecs, L=Lfull, diag=diag)


def local_orbitals(pos, nocc, seed, width=0.6):
    rng = np.random.default_rng(seed)
    n = len(pos)
    cen = rng.uniform(pos.min(), pos.max(), nocc)
    C = np.exp(-width * (pos[:, None] - cen[None, :]) ** 2) * rng.standard_normal((n, nocc))
    return C / np.linalg.norm(C, axis=0)


def store(mol):
    import oracle

    return oracle.CholStore(mol["n"], mol["sets"], list(zip(mol["vred"], mol["vecs"])))
