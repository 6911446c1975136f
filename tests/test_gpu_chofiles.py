"""A context loaded from the reference's Cholesky files (mjk_choinfo_read +
mjk_load_choinfo: CHORST, CHRED, CHVEC1, CHODIAG written in the reference's
layout by tests/_chofiles.py) builds bit-identically to one fed the same maps
and vectors from memory, for word- and block-addressed vector files, and its
LK-screened build (shells and diagonal taken from the files) equals the
screened build with shells and diagonal given in memory."""
import numpy as np
import pytest
import torch

from _chofiles import write_cholesky_files
from _lk_molecule import local_orbitals, molecule
from openmolcas_amd.fock import CholeskyFiles, JKContext

pytestmark = pytest.mark.gpu


def fold(a):
    f = 2.0 * a
    f[np.diag_indices(a.shape[0])] = np.diag(a)
    i, j = np.tril_indices(a.shape[0])
    return f[i, j]


@pytest.mark.parametrize("adrvec", [1, 2])
def test_context_from_files_matches_memory(gpu, tmp_path, adrvec):
    mol = molecule(natom=6, seed=101, nvec=60, nred=3)
    n = mol["n"]
    isoshl = np.repeat(np.arange(1, len(mol["nbassh"]) + 1), mol["nbassh"])
    maps, vecs, _ = write_cholesky_files(tmp_path, n, isoshl, mol["sets"], mol["vred"], mol["L"], mol["diag"],
                                         adrvec=adrvec, thrcom=1e-6)
    cf = CholeskyFiles(str(tmp_path / "CHORST"), str(tmp_path / "CHRED"), isoshl)
    fil = JKContext(n, cf.numcho, device=gpu)
    cf.load(fil, str(tmp_path / "CHVEC1"), str(tmp_path / "CHODIAG"))
    mem = JKContext(n, len(vecs), device=gpu)
    vred = np.asarray(mol["vred"])
    for r, m in enumerate(maps):
        mem.set_reduced_set(r, m)
    for r in range(len(maps)):
        idx = np.nonzero(vred == r)[0]
        mem.upload_vectors(r, int(idx[0]), np.stack([vecs[j] for j in idx]))
    C = local_orbitals(mol["pos"], 7, 102)
    dlt = torch.from_numpy(fold(C @ C.T))
    J0, K0 = mem.coulomb_exchange(dlt, [torch.from_numpy(C)])
    J1, K1 = fil.coulomb_exchange(dlt, [torch.from_numpy(C)])
    assert torch.equal(J0, J1) and torch.equal(K0, K1)
    # LK: the loader set the shells (consecutive here) and the CHODIAG diagonal
    mem.set_shells(mol["nbassh"])
    mem.set_lk_diag(mol["diag"], maxred=len(maps))
    for jk in (mem, fil):
        jk.set_lk_screening(True, thrcom=cf.thrcom)
    plt = fold(C @ C.T)
    f0, f1 = np.zeros((1, n * (n + 1) // 2)), np.zeros((1, n * (n + 1) // 2))
    mem.cho_lk_scf(1, f0, None, [0], [7], [C], plt, 0.5, nscreen=2, dmpk=0.2)
    fil.cho_lk_scf(1, f1, None, [0], [7], [C], plt, 0.5, nscreen=2, dmpk=0.2)
    assert np.array_equal(f0, f1)
    assert np.array_equal(mem.lk_stats(), fil.lk_stats())
    cf.close()
    mem.close()
    fil.close()


@pytest.mark.parametrize("adrvec", [1, 2])
def test_context_from_symmetric_files_matches_memory(gpu, tmp_path, adrvec):
    """Four irreps: mjk_load_choinfo_sym (one CHVEC file per irrep) fills a
    context whose J and K equal, bit for bit, those of a context given the
    same (irrep, reduced set) maps and vectors from memory; the exchange of
    the vectors of irreps 2-4 shows (they carry no Coulomb part for a totally
    symmetric density)."""
    import test_chofiles as tc
    from _chofiles import write_cholesky_files_sym

    sets, vec_sym, vec_set, vec_lt, diag = tc.sym_problem(seed=31, nvec=(9, 6, 5, 4))
    maps, vecs, _ = write_cholesky_files_sym(tmp_path, tc.NBAS_SYM, tc.ISOSHL_SYM, sets, vec_sym, vec_set,
                                             vec_lt, diag, adrvec=adrvec)
    cf = CholeskyFiles(str(tmp_path / "CHORST"), str(tmp_path / "CHRED"), tc.ISOSHL_SYM)
    n, nsym, maxred = cf.nbas, cf.nsym, cf.maxred
    fil = JKContext(n, cf.numcho, device=gpu)
    with pytest.raises(Exception):
        cf.load(fil, str(tmp_path / "CHVEC1"))        # one file cannot hold four irreps
    cf.load(fil, [str(tmp_path / f"CHVEC{s + 1}") for s in range(nsym)], str(tmp_path / "CHODIAG"))
    mem = JKContext(n, cf.numcho, device=gpu)
    for (s, r), m in maps.items():
        mem.set_reduced_set(s * maxred + r, m)
    j = 0
    for s in range(nsym):
        js = [J for J in range(len(vec_sym)) if vec_sym[J] == s]
        for r in range(maxred):
            rows = [vecs[s][k] for k, J in enumerate(js) if vec_set[J] == r]
            if rows:
                mem.upload_vectors(s * maxred + r, j, np.stack(rows))
                j += len(rows)
    assert j == cf.numcho
    # a totally symmetric density: orbitals inside irrep blocks
    rng = np.random.default_rng(32)
    C = np.zeros((n, 6))
    off = np.concatenate([[0], np.cumsum(tc.NBAS_SYM)])
    for k, s in enumerate([0, 0, 1, 2, 2, 3]):
        C[off[s]:off[s + 1], k] = rng.standard_normal(off[s + 1] - off[s])
    dlt = torch.from_numpy(fold(C @ C.T))
    J0, K0 = mem.coulomb_exchange(dlt, [torch.from_numpy(C)])
    J1, K1 = fil.coulomb_exchange(dlt, [torch.from_numpy(C)])
    assert torch.equal(J0, J1) and torch.equal(K0, K1)
    # against the dense formula over all vectors, every irrep included
    L = np.stack([vec_lt[J] for J in range(len(vec_lt))])
    i, jj = np.tril_indices(n)
    Ls = np.zeros((len(L), n, n))
    Ls[:, i, jj] = L
    Ls[:, jj, i] = L
    Jd = np.einsum("Jab,Jcd,cd->ab", Ls, Ls, C @ C.T)[i, jj]
    X = np.einsum("Jab,bk->Jak", Ls, C)
    Kd = np.einsum("Jak,Jbk->ab", X, X)[i, jj]
    assert np.max(np.abs(J1.cpu().numpy().reshape(-1) - Jd)) <= 1e-10 * max(1.0, np.max(np.abs(Jd)))
    assert np.max(np.abs(K1.cpu().numpy().reshape(-1) - Kd)) <= 1e-10 * max(1.0, np.max(np.abs(Kd)))
    assert np.max(np.abs(Kd)) > 0 and np.max(np.abs(np.einsum("Jab,bk->Jak", Ls[9:], C))) > 0
    cf.close()
    mem.close()
    fil.close()
