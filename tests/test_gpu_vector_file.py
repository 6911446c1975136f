"""Vectors read straight from a vector file in the reference's on-disk layout
(mjk_upload_vectors_file).

The file is written here the way the reference stores CHVEC<iSym>: FP64
words, vector J of reduced set iRed at word address InfVec(J,3) holding
nnBstR(iSym,iRed) words (src/cholesky_util/cho_vecrd1.F90:93-147, word
addressing src/io_util/ddafile.F90:35-36).  Two reduced sets (all pairs, and a
screened subset), vectors at running addresses with gaps.  The build from the
file must be bit-identical to the build from the same vectors uploaded from
memory, also when the unit is split into pieces of a fixed byte size the way
MOLCAS_MAXFILESIZE splits it (src/io_util/mpdafile.F90:92-140: pieces named
path, path'1'.., vectors straddling pieces); a missing file gives MJK_ERR_IO
(4) and an address past the end of the file MJK_ERR_SHORT_READ (77)."""
import numpy as np
import pytest
import torch

from openmolcas_amd import MjkError
from openmolcas_amd.fock import JKContext
from openmolcas_amd.synthetic import chol_vectors_np, fold, npair, orbitals_np

pytestmark = pytest.mark.gpu


def layout(n, M0, M1, seed):
    rng = np.random.default_rng(seed)
    npr = npair(n)
    rs1 = np.sort(rng.choice(npr, size=npr // 3, replace=False)).astype(np.int64)
    L = chol_vectors_np(n, M0 + M1, seed=seed)
    L0 = L[:M0]
    L1 = L[M0:, rs1]
    return rs1, L0, L1


def write_vector_file(path, blocks, start=5, gap=3):
    """blocks: list of (jnum, nrs) arrays; returns the word address per vector."""
    words = [np.zeros(start)]
    iadr, at = [], start
    for blk in blocks:
        for v in blk:
            iadr.append(at)
            words.append(v)
            at += v.size
        words.append(np.full(gap, np.nan))  # never read
        at += gap
    np.concatenate(words).astype("<f8").tofile(path)
    return np.array(iadr, dtype=np.int64)


def test_vectors_from_file_match_memory(gpu, tmp_path):
    n, M0, M1 = 40, 10, 7
    npr = npair(n)
    rs1, L0, L1 = layout(n, M0, M1, seed=3)
    path = str(tmp_path / "CHVEC1")
    iadr = write_vector_file(path, [L0, L1])
    C = torch.from_numpy(orbitals_np(n, 6, seed=4))
    A = np.random.default_rng(5).standard_normal((n, n))
    dlt = torch.from_numpy(fold((A + A.T) / (2.0 * n)))

    mem = JKContext(n, M0 + M1, device=gpu)
    mem.set_reduced_set(0, np.arange(npr))
    mem.set_reduced_set(1, rs1)
    mem.upload_vectors(0, 0, L0)
    mem.upload_vectors(1, M0, L1)

    fil = JKContext(n, M0 + M1, device=gpu)
    fil.set_reduced_set(0, np.arange(npr))
    fil.set_reduced_set(1, rs1)
    fil.upload_vectors_file(path, 0, 0, iadr[:M0])
    fil.upload_vectors_file(path, 1, M0, iadr[M0:])

    J0, K0 = mem.coulomb_exchange(dlt, [C])
    J1, K1 = fil.coulomb_exchange(dlt, [C])
    assert torch.equal(J0, J1) and torch.equal(K0, K1)
    mem.close()
    fil.close()


def split_file(path, piece):
    """Splits a logical unit into pieces of `piece` bytes: path, path + '1', ..."""
    data = open(path, "rb").read()
    names = []
    for k in range(0, (len(data) + piece - 1) // piece):
        name = path if k == 0 else path + (chr(ord("0") + k) if k <= 9 else chr(ord("A") + k - 10))
        names.append(name)
        with open(name, "wb") as f:
            f.write(data[k * piece:(k + 1) * piece])
    return names


def test_vectors_from_split_file(gpu, tmp_path):
    n, M0, M1 = 40, 10, 7
    npr = npair(n)
    rs1, L0, L1 = layout(n, M0, M1, seed=8)
    path = str(tmp_path / "CHVEC1")
    iadr = write_vector_file(path, [L0, L1])
    piece = 8000  # a multiple of 8 bytes, as MaxFileSize * 10^6 is
    names = split_file(path, piece)
    assert 10 <= len(names) <= 20
    C = torch.from_numpy(orbitals_np(n, 5, seed=9))
    A = np.random.default_rng(10).standard_normal((n, n))
    dlt = torch.from_numpy(fold((A + A.T) / (2.0 * n)))
    mem = JKContext(n, M0 + M1, device=gpu)
    mem.set_reduced_set(0, np.arange(npr))
    mem.set_reduced_set(1, rs1)
    mem.upload_vectors(0, 0, L0)
    mem.upload_vectors(1, M0, L1)
    fil = JKContext(n, M0 + M1, device=gpu)
    fil.set_reduced_set(0, np.arange(npr))
    fil.set_reduced_set(1, rs1)
    fil.upload_vectors_file(path, 0, 0, iadr[:M0], split_bytes=piece)
    fil.upload_vectors_file(path, 1, M0, iadr[M0:], split_bytes=piece)
    J0, K0 = mem.coulomb_exchange(dlt, [C])
    J1, K1 = fil.coulomb_exchange(dlt, [C])
    assert torch.equal(J0, J1) and torch.equal(K0, K1)
    mem.close()
    fil.close()


def test_vector_file_errors(gpu, tmp_path):
    n, M0 = 12, 3
    L0 = chol_vectors_np(n, M0, seed=6)
    path = str(tmp_path / "CHVEC1")
    iadr = write_vector_file(path, [L0])
    jk = JKContext(n, M0, device=gpu)
    jk.set_reduced_set(0, np.arange(npair(n)))
    with pytest.raises(MjkError) as e:
        jk.upload_vectors_file(str(tmp_path / "missing"), 0, 0, iadr)
    assert e.value.rc == 4
    with pytest.raises(MjkError) as e:
        jk.upload_vectors_file(path, 0, 0, iadr + 10)  # the last vector runs past the end
    assert e.value.rc == 77
    with pytest.raises(MjkError) as e:
        jk.upload_vectors_file(path, 1, 0, iadr)  # reduced set never declared
    assert e.value.rc == 3
    jk.upload_vectors_file(path, 0, 0, iadr)
    C = torch.from_numpy(orbitals_np(n, 2, seed=7))
    K = jk.exchange([C])
    ref = JKContext(n, M0, device=gpu)
    ref.upload_packed(L0)
    assert torch.equal(K, ref.exchange([C]))
    jk.close()
    ref.close()


def write_block_file(path, blocks, mbl=512):
    """CHO_ADRVEC = 2 layout: the file is opened with DANAME_MF, so a disk
    address counts blocks of MBL = 512 bytes (src/io_util/ddafile.F90:36,
    src/io_util/fast_io.F90:81) and every write ends on the next block
    boundary (ddafile.F90:42).  Returns the block address per vector."""
    iadr, chunks, at = [], [], 2  # the first two blocks hold something else
    chunks.append(np.full(at * mbl // 8, np.nan))
    for blk in blocks:
        for v in blk:
            iadr.append(at)
            nblk = -(-v.size * 8 // mbl)
            pad = np.full(nblk * mbl // 8 - v.size, np.nan)  # never read
            chunks += [v, pad]
            at += nblk
    np.concatenate(chunks).astype("<f8").tofile(path)
    return np.array(iadr, dtype=np.int64)


def test_block_addressed_vector_file(gpu, tmp_path):
    n, M0, M1 = 33, 6, 5
    npr = npair(n)
    rs1, L0, L1 = layout(n, M0, M1, seed=13)
    path = str(tmp_path / "CHVEC1")
    iadr = write_block_file(path, [L0, L1])
    C = torch.from_numpy(orbitals_np(n, 5, seed=14))
    dlt = torch.from_numpy(fold(np.eye(n)))
    mem = JKContext(n, M0 + M1, device=gpu)
    fil = JKContext(n, M0 + M1, device=gpu)
    for jk in (mem, fil):
        jk.set_reduced_set(0, np.arange(npr))
        jk.set_reduced_set(1, rs1)
    mem.upload_vectors(0, 0, L0)
    mem.upload_vectors(1, M0, L1)
    fil.upload_vectors_file(path, 0, 0, iadr[:M0], mbl=512)
    fil.upload_vectors_file(path, 1, M0, iadr[M0:], mbl=512)
    J0, K0 = mem.coulomb_exchange(dlt, [C])
    J1, K1 = fil.coulomb_exchange(dlt, [C])
    assert torch.equal(J0, J1) and torch.equal(K0, K1)
    with pytest.raises(MjkError) as e:  # only word (8) and block (512) addressing exist
        fil.upload_vectors_file(path, 0, 0, iadr[:M0], mbl=7)
    assert e.value.rc == 1
    mem.close()
    fil.close()
