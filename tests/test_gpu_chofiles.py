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
