"""Python mirror of OpenMolcas's Cholesky J/K Fock-build interface.

``JKContext`` owns one library context: the Cholesky vectors resident in HBM
(the role of Cho_X_Init's vector buffer, src/cholesky_util/cho_x_init.F90:55)
and the device workspaces.  Two layers are exposed:

* ``cho_fscf`` / ``cho_fmcscf`` take host (numpy) arrays with the reference's
  argument meaning and storage (src/fock_util/cho_fscf.F90:14,
  src/fock_util/cho_fmcscf.F90:14); FLT is updated in place (``+=``) and a
  nonzero return code raises ``MjkError``, where the reference would Abend.
* ``coulomb`` / ``exchange`` / ``fock_update`` / ``energy`` / ``puvx`` work on
  torch device tensors, queued on torch's current stream.

Every call goes through the C ABI of ``libmolcas_jk.so``; there is no CPU path.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np
import torch

from . import _lib
from ._lib import MjkError, PREC_FP64

_PREC = {"fp64": 0, "fp32": 1, "bf16x3": 2, "ladder": 3}


def precision_code(p) -> int:
    if isinstance(p, str):
        return _PREC[p]
    return int(p)


def _ptr(t) -> ctypes.c_void_p:
    if isinstance(t, torch.Tensor):
        return ctypes.c_void_p(t.data_ptr())
    if isinstance(t, np.ndarray):
        return ctypes.c_void_p(t.ctypes.data)
    if t is None:
        return ctypes.c_void_p(0)
    raise TypeError(type(t))


def _dev_f64(t: torch.Tensor, device) -> torch.Tensor:
    if not isinstance(t, torch.Tensor):
        t = torch.as_tensor(np.asarray(t))
    return t.to(device=device, dtype=torch.float64).contiguous()


def _ld(t: torch.Tensor) -> int:
    """Row stride of a contiguous 2-D tensor (stride(0) is arbitrary when the
    leading dimension is 1)."""
    if t.dim() != 2 or not t.is_contiguous():
        raise ValueError("expected a contiguous 2-D tensor")
    return t.shape[1]


def _host_f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def _rows_inout(a, n: int, npair: int, what: str):
    """Per-density host arrays updated in place: a (n, npair) C-contiguous
    float64 array (its rows) or a list of n separately allocated 1-D float64
    arrays.  Returns the list of writable arrays."""
    rows = list(a) if isinstance(a, (list, tuple)) else [a[d] for d in range(n)]
    if len(rows) < n:
        raise ValueError(f"{what}: expected {n} per-density arrays")
    for r in rows[:n]:
        if not isinstance(r, np.ndarray) or r.dtype != np.float64 or not r.flags.c_contiguous \
                or r.size < npair:
            raise TypeError(f"{what} must be C-contiguous float64 arrays of >= {npair} values "
                            "(updated in place)")
    return rows[:n]


def _ptr_array(arrs):
    """C array of pointers to the given host arrays (None -> NULL)."""
    return (ctypes.c_void_p * max(1, len(arrs)))(
        *[a.ctypes.data if a is not None else None for a in arrs])


class JKContext:
    """One device context for nbas basis functions and nchol vectors (nSym=1)."""

    def __init__(self, nbas: int, nchol: int, device=None, share: "JKContext" = None):
        if not torch.cuda.is_available():
            raise RuntimeError("JKContext needs a HIP device (MI355X); there is no CPU fallback")
        self.lib = _lib.load()
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.nbas = int(nbas)
        self.nchol = int(nchol)
        self.npair = self.nbas * (self.nbas + 1) // 2
        h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            if share is None:
                rc = self.lib.mjk_create(ctypes.byref(h), self.nbas, self.nchol)
            else:
                rc = self.lib.mjk_create_shared(ctypes.byref(h), self.nbas, self.nchol, share._h)
        if rc != 0:
            raise MjkError(rc, "mjk_create failed")
        self._h = h
        self._share = share  # keeps the workspace owner alive
        if share is None:
            self.bind_stream(torch.cuda.current_stream(self.device))
        else:
            self.stream = share.stream

    # ------------------------------------------------------------ plumbing
    def _check(self, rc: int):
        if rc != 0:
            msg = self.lib.mjk_last_error(self._h)
            raise MjkError(rc, msg.decode() if msg else "")

    def bind_stream(self, stream: torch.cuda.Stream):
        """Queues all device work on ``stream`` (torch's current stream by default)."""
        self._check(self.lib.mjk_set_stream(self._h, ctypes.c_void_p(stream.cuda_stream)))
        self.stream = stream

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self.lib.mjk_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_ladder(self, j_fp32: int, j_bf16: int) -> None:
        """Split points of the "ladder" precision (FP64 / FP32 / split-BF16)."""
        self._check(self.lib.mjk_set_ladder(self._h, int(j_fp32), int(j_bf16)))

    def set_x_budget(self, nbytes: int):
        self._check(self.lib.mjk_set_x_budget(self._h, int(nbytes)))

    def sync(self):
        self._check(self.lib.mjk_sync(self._h))

    # ------------------------------------------------------------ vectors
    def set_reduced_set(self, jred: int, rs_to_lt) -> None:
        """Reduced set jred: rs_to_lt[r] = 0-based LT index of pair r."""
        m = np.ascontiguousarray(rs_to_lt, dtype=np.int64)
        self._check(self.lib.mjk_set_reduced_set(self._h, int(jred), len(m), _ptr(m)))

    def upload_vectors(self, jred: int, jvec: int, lrs) -> None:
        """Vectors jvec.. of reduced set jred; lrs is (jnum, nrs), one vector per
        row (the memory of the reference's Lrs(nRS, JNUM)).  numpy -> host copy,
        torch CUDA tensor -> device copy."""
        if isinstance(lrs, torch.Tensor) and lrs.is_cuda:
            t = lrs.to(torch.float64).contiguous()
            self._check(self.lib.mjk_upload_vectors(self._h, int(jred), int(jvec), t.shape[0], _ptr(t), 1))
        else:
            a = _host_f64(lrs)
            self._check(self.lib.mjk_upload_vectors(self._h, int(jred), int(jvec), a.shape[0], _ptr(a), 0))

    def upload_vectors_file(self, path: str, jred: int, jvec: int, iadr, split_bytes: int = 0,
                            mbl: int = 8) -> None:
        """Vectors jvec.. of reduced set jred read from the reference's vector
        file CHVEC<iSym>: iadr[i] is the word address (InfVec(J,3,iSym)) of
        vector jvec+i, each holding the reduced set's nrs doubles.  split_bytes:
        piece size of a unit split by MOLCAS_MAXFILESIZE (0: one file); mbl:
        bytes per disk address unit (8 word-addressed, 512 block-addressed)."""
        a = np.ascontiguousarray(np.asarray(iadr, dtype=np.int64))
        self._check(self.lib.mjk_upload_vectors_file(self._h, os.fsencode(path), int(jred), int(jvec),
                                                     a.shape[0], _ptr(a), int(mbl), int(split_bytes)))

    def upload_packed(self, L) -> None:
        """All vectors on one reduced set covering every LT pair: L is (nchol, npair)."""
        if L.shape != (self.nchol, self.npair):
            raise ValueError(f"expected L of shape {(self.nchol, self.npair)}, got {tuple(L.shape)}")
        self.set_reduced_set(0, np.arange(self.npair, dtype=np.int64))
        self.upload_vectors(0, 0, L)

    # ------------------------------------------------------------ drop-ins (host)
    def cho_fscf(self, nden: int, flt, nforb, niorb, porb, dlt: np.ndarray,
                 exfac: float, precision="fp64") -> int:
        """CHO_FSCF(rc,nDen,FLT,nForb,nIorb,Porb,DLT,ExFac), nSym = 1.

        flt: (nden, npair) float64 array or a list of nden separate arrays,
        updated in place (FLT(d)%A0).  porb: list of nden arrays (nbas,
        nforb[d]+niorb[d]) = Porb(d)%SB(1)%A2(nOcc,nBas) memory.  dlt: folded
        LT density (only DLT(1) enters Coulomb, as in the reference)."""
        nd = int(nden) if nden in (1, 2) else 1  # an invalid nDen is rejected by the library
        fl = _rows_inout(flt, nd, self.npair, "flt")
        nf = np.ascontiguousarray(nforb, dtype=np.int32)
        ni = np.ascontiguousarray(niorb, dtype=np.int32)
        d = _host_f64(dlt).reshape(-1)[: self.npair].copy()
        pc = [_host_f64(porb[k]) for k in range(nd)] if nden in (1, 2) else [None] * nd
        rc = self.lib.mjk_cho_fscf(self._h, int(nden), _ptr_array(fl), _ptr(nf), _ptr(ni),
                                   _ptr_array(pc), _ptr(d), float(exfac), precision_code(precision))
        self._check(rc)
        return rc

    def cho_fscf_sym(self, nbas, nden: int, flt: np.ndarray, nforb, niorb, porb,
                     dlt: np.ndarray, exfac: float, precision="fp64") -> int:
        """CHO_FSCF with point-group symmetry (nSym = len(nbas) irreps, basis
        ordered irrep by irrep).

        flt: (nden, sum_s nbas[s](nbas[s]+1)/2) float64, the irrep LT blocks
        back to back, updated in place; dlt: the same layout (folded).  nforb,
        niorb: (nden, nsym) counts.  porb[d][s]: (nbas[s], nocc(s)) =
        Porb(d)%SB(s)%A2(nOcc,nBas) memory.  The vectors (every JSYM) must be
        in the context already, with reduced-set maps into the C1 triangle."""
        nsym = len(nbas)
        ntri = sum(int(b) * (int(b) + 1) // 2 for b in nbas)
        fl = _rows_inout(flt, nden, ntri, "flt")
        nb = np.ascontiguousarray(nbas, dtype=np.int32)
        nf = np.zeros((max(nden, 1), 8), dtype=np.int32)
        ni = np.zeros((max(nden, 1), 8), dtype=np.int32)
        nf[:nden, :nsym] = np.asarray(nforb, dtype=np.int32).reshape(nden, nsym)
        ni[:nden, :nsym] = np.asarray(niorb, dtype=np.int32).reshape(nden, nsym)
        pc = []  # Porb(d)%A0: the irrep blocks of density d back to back
        for dd in range(nden):
            parts = [_host_f64(porb[dd][k]).reshape(-1) for k in range(nsym)]
            pc.append(np.concatenate(parts) if parts else np.zeros(1))
        d = _host_f64(dlt).reshape(-1)
        rc = self.lib.mjk_cho_fscf_sym(self._h, int(nsym), _ptr(nb), int(nden), _ptr_array(fl),
                                       _ptr(nf), _ptr(ni), _ptr_array(pc), _ptr(d), float(exfac),
                                       precision_code(precision))
        self._check(rc)
        return rc

    def cho_lk_scf_sym(self, nbas, nden: int, flt: np.ndarray, nforb, niorb, porb,
                       plt: np.ndarray, factxi: float, nscreen: int = 10, dmpk: float = 0.1,
                       dfmat: float = 0.0, klt=None, nbassh=None, precision="fp64") -> int:
        """CHO_LK_SCF with point-group symmetry (storage as cho_fscf_sym),
        unscreened.

        porb[d][s]: (nbas[s], >= nocc(s)) AO x orbital matrices (the columns of
        Porb(d)%SB(s)%A2(nBas,nBas)); plt: folded total density in irrep LT
        blocks; flt updated in place.  klt: None or the same form as flt,
        accumulated with -FactXI K in each irrep's shell-pair-blocked storage;
        nbassh: (nsym, nshell) shell sizes nBasSh (None: one shell per irrep)."""
        nsym = len(nbas)
        ntri = sum(int(b) * (int(b) + 1) // 2 for b in nbas)
        fl = _rows_inout(flt, nden, ntri, "flt")
        kl = _rows_inout(klt, nden, ntri, "klt") if klt is not None else None
        nb = np.ascontiguousarray(nbas, dtype=np.int32)
        nf = np.zeros((max(nden, 1), 8), dtype=np.int32)
        ni = np.zeros((max(nden, 1), 8), dtype=np.int32)
        nf[:nden, :nsym] = np.asarray(nforb, dtype=np.int32).reshape(nden, nsym)
        ni[:nden, :nsym] = np.asarray(niorb, dtype=np.int32).reshape(nden, nsym)
        pc = []  # Porb(d)%A0: square irrep blocks, AO index fastest
        for d in range(nden):
            parts = []
            for k in range(nsym):
                c = _host_f64(porb[d][k])
                sq = np.zeros((nbas[k], nbas[k]))
                sq[:, :c.shape[1]] = c
                parts.append(sq.T.reshape(-1))
            pc.append(np.concatenate(parts) if parts else np.zeros(1))
        p = _host_f64(plt).reshape(-1)
        nshell = 0
        sh = np.zeros(1, dtype=np.int32)
        if nbassh is not None:
            shm = np.asarray(nbassh, dtype=np.int32).reshape(nsym, -1)
            nshell = shm.shape[1]
            sh = np.ascontiguousarray(shm.T).reshape(-1)  # nBasSh(iSym, iSh), iSym fastest
        rc = self.lib.mjk_cho_lk_scf_sym(self._h, int(nsym), _ptr(nb), int(nden), _ptr_array(fl),
                                         _ptr_array(kl) if kl is not None else None, _ptr(nf), _ptr(ni),
                                         _ptr_array(pc), _ptr(p), float(factxi), int(nscreen), float(dmpk),
                                         float(dfmat), int(nshell), _ptr(sh), precision_code(precision))
        self._check(rc)
        return rc

    def cho_fock_dft_red(self, dlt: np.ndarray, flt: np.ndarray) -> int:
        """CHO_FOCK_DFT_RED(irc,DLT,FLT): Coulomb-only Fock matrix, flt (npair,)
        overwritten with J[DLT] (the reference zeroes FLT first)."""
        if flt.dtype != np.float64 or not flt.flags.c_contiguous or flt.size < self.npair:
            raise TypeError("flt must be a C-contiguous float64 array of npair values")
        d = _host_f64(dlt).reshape(-1)[: self.npair].copy()
        rc = self.lib.mjk_cho_fock_dft_red(self._h, _ptr(d), _ptr(flt))
        self._check(rc)
        return rc

    def cho_lk_scf(self, nden: int, flt, klt, nforb, niorb, porb, plt: np.ndarray,
                   factxi: float, nscreen: int = 10, dmpk: float = 0.1, dfmat: float = 0.0,
                   nbassh=None, precision="fp64") -> int:
        """CHO_LK_SCF(rc,nDen,FLT,KLT,nForb,nIorb,Porb,PLT,FactXI,nScreen,dmpk,
        dFmat), nSym = 1 (screened after set_lk_screening, else the dmpk -> 0
        result).

        flt: (nden, npair) array or list of nden arrays, updated in place.
        klt: None, or the same form, accumulated in the shell-pair-blocked
        storage of the shells nbassh.  porb: list of nden (nbas, >=
        nforb[d]+niorb[d]) AO x orbital matrices C[b, k] (the MSQ columns).
        plt: folded total density."""
        nd = int(nden) if nden in (1, 2) else 1  # an invalid nDen is rejected by the library
        fl = _rows_inout(flt, nd, self.npair, "flt")
        kl = _rows_inout(klt, nd, self.npair, "klt") if klt is not None else None
        nf = np.ascontiguousarray(nforb, dtype=np.int32)
        ni = np.ascontiguousarray(niorb, dtype=np.int32)
        n = self.nbas
        pc = [None] * nd
        if nden in (1, 2):
            for d in range(nden):
                c = _host_f64(porb[d])
                sq = np.zeros((n, n))
                sq[:, :c.shape[1]] = c
                pc[d] = np.ascontiguousarray(sq.T)  # A2(nBas,nBas), AO index fastest
        p = _host_f64(plt).reshape(-1)[: self.npair].copy()
        sh = np.ascontiguousarray(nbassh if nbassh is not None else [], dtype=np.int32)
        rc = self.lib.mjk_cho_lk_scf(self._h, int(nden), _ptr_array(fl),
                                     _ptr_array(kl) if kl is not None else None, _ptr(nf), _ptr(ni),
                                     _ptr_array(pc), _ptr(p), float(factxi), int(nscreen),
                                     float(dmpk), float(dfmat), len(sh), _ptr(sh) if len(sh) else None,
                                     precision_code(precision))
        self._check(rc)
        return rc

    def cho_lk_casscf(self, dlt: np.ndarray, flt: np.ndarray, msq, w_pwxy, factxi: float,
                      nfiorb: int, naorb: int, nchm: int, ash1, ash2, do_active: bool, cmo,
                      nfro: int, norb: int, nscreen: int = 10, dmpk: float = 0.1,
                      dfmat: float = 0.0, exfac: float = 1.0, precision="fp64") -> int:
        """CHO_LK_CASSCF(DLT,FLT,MSQ,W_PWXY,FactXI,nFIorb,nAorb,nChM,Ash,DoActive,
        nScreen,dmpk,dFmat,CMO,ExFac), nSym = 1, without screening.  flt (2, npair)
        updated in place; msq (nbas, >= nfiorb) and ash2 (nbas, >= nchm) are AO x
        vector matrices; ash1 (nbas, naorb) the active MOs; cmo (nbas, nbas).
        flt: (2, npair) or a list [FI, FA]; dlt: (2, npair) or [DI, DA]."""
        fl = _rows_inout(flt, 2 if do_active else 1, self.npair, "flt")
        n = self.nbas

        def square(a, k):  # A2(nBas,nBas), AO index fastest
            sq = np.zeros((n, n))
            if a is not None and k > 0:
                sq[:, :k] = _host_f64(a)[:, :k]
            return np.ascontiguousarray(sq.T)

        m = square(msq, nfiorb)
        a2 = square(ash2, nchm)
        a1 = _host_f64(ash1) if ash1 is not None and naorb > 0 else np.zeros(1)
        cm = _host_f64(np.asarray(cmo).T) if cmo is not None else np.zeros(1)
        dl = [_host_f64(x).reshape(-1) for x in (list(dlt) if isinstance(dlt, (list, tuple))
                                                 else np.asarray(dlt).reshape(-1, self.npair))]
        rc = self.lib.mjk_cho_lk_casscf(self._h, _ptr_array(dl), _ptr_array(fl), _ptr(m), _ptr(w_pwxy),
                                        float(factxi), int(nfiorb), int(naorb), int(nchm),
                                        _ptr(a1), _ptr(a2), int(bool(do_active)), int(nscreen),
                                        float(dmpk), float(dfmat), _ptr(cm), int(nfro), int(norb),
                                        float(exfac), precision_code(precision))
        self._check(rc)
        return rc

    def cho_fmcscf(self, flt: np.ndarray, nforb: int, niorb: int, naorb: int, factxi: float,
                   dlt: np.ndarray, do_active: bool, porb1, nchm: int, porb2, porb3,
                   w_pwxy, cmo, nfro: int, norb: int, exfac: float = 1.0,
                   precision="fp64") -> int:
        """CHO_FMCSCF(rc,FLT,nForb,nIorb,nAorb,FactXI,DLT,DoActive,POrb,nChM,
        W_PWXY,CMO,ExFac), nSym = 1.  flt (2, npair) updated in place; w_pwxy
        (norb*naorb*naorb(naorb+1)/2,) filled; cmo (nbas, nbas) with cmo[a, p] =
        CMO(a,p).  flt: (2, npair) or a list [FI, FA] of separate arrays;
        dlt: (2, npair) or [DI, DA]."""
        fl = _rows_inout(flt, 2 if do_active else 1, self.npair, "flt")
        p1 = _host_f64(porb1) if porb1 is not None else np.zeros(1)
        p2 = _host_f64(porb2) if porb2 is not None else np.zeros(1)
        p3 = _host_f64(porb3) if porb3 is not None else np.zeros(1)
        cm = _host_f64(np.asarray(cmo).T) if cmo is not None else np.zeros(1)
        dl = [_host_f64(x).reshape(-1) for x in (list(dlt) if isinstance(dlt, (list, tuple))
                                                 else np.asarray(dlt).reshape(-1, self.npair))]
        if w_pwxy is not None and (w_pwxy.dtype != np.float64 or not w_pwxy.flags.c_contiguous):
            raise TypeError("w_pwxy must be a C-contiguous float64 array")
        rc = self.lib.mjk_cho_fmcscf(self._h, _ptr_array(fl), int(nforb), int(niorb), int(naorb),
                                     float(factxi), _ptr_array(dl), int(bool(do_active)), _ptr(p1),
                                     int(nchm), _ptr(p2), _ptr(p3), _ptr(w_pwxy), _ptr(cm),
                                     int(nfro), int(norb), float(exfac), precision_code(precision))
        self._check(rc)
        return rc

    def cho_fmcscf_sym(self, nbas, flt, nforb, niorb, naorb, factxi: float, dlt,
                       do_active: bool, porb1, nchm, porb2, porb3, w_pwxy, cmo, nfro, norb,
                       exfac: float = 1.0, precision="fp64") -> int:
        """CHO_FMCSCF with point-group symmetry (nSym = len(nbas) irreps, basis
        ordered irrep by irrep).

        flt, dlt: (2, ntri) or [FI, FA] / [DI, DA], the irrep LT blocks back to
        back (flt updated in place).  nforb, niorb, naorb, nchm, nfro, norb:
        per-irrep counts.  porb1[s], porb2[s], porb3[s]: (nbas[s], nk(s)) AO x
        orbital matrices (POrb(i)%SB(s)%A2(nk,nBas) memory); cmo[s]: (nbas[s],
        nbas[s]) with cmo[s][a, p] = CMO%SB(s)%A2(a,p).  w_pwxy: filled in the
        symmetry-blocked (pw|xy) order of CHO_eval_waxy."""
        nsym = len(nbas)
        ntri = sum(int(b) * (int(b) + 1) // 2 for b in nbas)
        nd = 2 if do_active else 1
        fl = _rows_inout(flt, nd, ntri, "flt")
        dl = [_host_f64(x).reshape(-1) for x in (list(dlt) if isinstance(dlt, (list, tuple))
                                                 else np.asarray(dlt).reshape(-1, ntri))]
        counts = [np.ascontiguousarray(np.asarray(v, dtype=np.int32).reshape(nsym))
                  for v in (nforb, niorb, naorb, nchm, nfro, norb)]

        def blocks(mats, ks):
            parts = [_host_f64(mats[s])[:, :ks[s]].copy().reshape(-1)
                     for s in range(nsym) if mats is not None and ks[s] > 0]
            return np.concatenate(parts) if parts else np.zeros(1)

        p1 = blocks(porb1, counts[0] + counts[1])
        p2 = blocks(porb2, counts[3]) if do_active else np.zeros(1)
        p3 = blocks(porb3, counts[2])
        cm = (np.concatenate([_host_f64(np.asarray(cmo[s]).T).reshape(-1) for s in range(nsym)])
              if cmo is not None else np.zeros(1))  # CMO%SB(s)%A2, column-major
        if w_pwxy is not None and (w_pwxy.dtype != np.float64 or not w_pwxy.flags.c_contiguous):
            raise TypeError("w_pwxy must be a C-contiguous float64 array")
        nb = np.ascontiguousarray(nbas, dtype=np.int32)
        rc = self.lib.mjk_cho_fmcscf_sym(self._h, int(nsym), _ptr(nb), _ptr_array(fl),
                                         _ptr(counts[0]), _ptr(counts[1]), _ptr(counts[2]),
                                         float(factxi), _ptr_array(dl), int(bool(do_active)),
                                         _ptr(p1), _ptr(counts[3]), _ptr(p2), _ptr(p3),
                                         _ptr(w_pwxy), _ptr(cm), _ptr(counts[4]), _ptr(counts[5]),
                                         float(exfac), precision_code(precision))
        self._check(rc)
        return rc

    def cho_lk_casscf_sym(self, nbas, dlt, flt, msq, w_pwxy, factxi: float, nchi, naorb, nchm,
                          ash1, ash2, do_active: bool, cmo, nfro, norb, nscreen: int = 10,
                          dmpk: float = 0.1, dfmat: float = 0.0, exfac: float = 1.0,
                          precision="fp64") -> int:
        """CHO_LK_CASSCF with point-group symmetry, unscreened (storage as
        cho_fmcscf_sym).  msq[s] (nbas[s], >= nchi[s]) and ash2[s] (nbas[s], >=
        nchm[s]) are AO x vector matrices (the MSQ / CVa(2) columns), ash1[s]
        (nbas[s], naorb[s]) the active MOs, cmo[s] (nbas[s], nbas[s])."""
        nsym = len(nbas)
        ntri = sum(int(b) * (int(b) + 1) // 2 for b in nbas)
        fl = _rows_inout(flt, 2 if do_active else 1, ntri, "flt")
        dl = [_host_f64(x).reshape(-1) for x in (list(dlt) if isinstance(dlt, (list, tuple))
                                                 else np.asarray(dlt).reshape(-1, ntri))]
        counts = [np.ascontiguousarray(np.asarray(v, dtype=np.int32).reshape(nsym))
                  for v in (nchi, naorb, nchm, nfro, norb)]

        def squares(mats, ks):  # A2(nBas(s),nBas(s)) blocks, AO index fastest
            parts = []
            for s in range(nsym):
                sq = np.zeros((nbas[s], nbas[s]))
                if mats is not None and ks[s] > 0:
                    sq[:, :ks[s]] = _host_f64(mats[s])[:, :ks[s]]
                parts.append(sq.T.reshape(-1))
            return np.concatenate(parts)

        m = squares(msq, counts[0])
        a2 = squares(ash2 if do_active else None, counts[2])
        parts = [_host_f64(ash1[s])[:, :counts[1][s]].copy().reshape(-1)
                 for s in range(nsym) if ash1 is not None and counts[1][s] > 0]
        a1 = np.concatenate(parts) if parts else np.zeros(1)
        cm = (np.concatenate([_host_f64(np.asarray(cmo[s]).T).reshape(-1) for s in range(nsym)])
              if cmo is not None else np.zeros(1))
        if w_pwxy is not None and (w_pwxy.dtype != np.float64 or not w_pwxy.flags.c_contiguous):
            raise TypeError("w_pwxy must be a C-contiguous float64 array")
        nb = np.ascontiguousarray(nbas, dtype=np.int32)
        rc = self.lib.mjk_cho_lk_casscf_sym(self._h, int(nsym), _ptr(nb), _ptr_array(dl), _ptr_array(fl),
                                            _ptr(m), _ptr(w_pwxy), float(factxi), _ptr(counts[0]),
                                            _ptr(counts[1]), _ptr(counts[2]), _ptr(a1), _ptr(a2),
                                            int(bool(do_active)), int(nscreen), float(dmpk), float(dfmat),
                                            _ptr(cm), _ptr(counts[3]), _ptr(counts[4]), float(exfac),
                                            precision_code(precision))
        self._check(rc)
        return rc

    # ------------------------------------------------------------ device engine
    def coulomb(self, dlt: torch.Tensor, want_v: bool = False):
        """J for folded LT densities dlt (nd, npair) on the device."""
        d = _dev_f64(dlt, self.device)
        if d.dim() == 1:
            d = d.unsqueeze(0)
        nd = d.shape[0]
        j = torch.empty((nd, self.npair), dtype=torch.float64, device=self.device)
        v = torch.empty((self.nchol, nd), dtype=torch.float64, device=self.device) if want_v else None
        self._check(self.lib.mjk_coulomb(self._h, nd, _ptr(d), _ld(d), _ptr(j), _ld(j), _ptr(v)))
        return (j, v) if want_v else j

    def _z_out(self, cs, z):
        if len(cs) != 1 or cs[0].dim() != 2 or not 1 <= cs[0].shape[1] <= 256:
            raise ValueError("the Z output needs exactly one orbital matrix of 1..256 columns")
        o = cs[0].shape[1]
        if (z.dtype != torch.float64 or z.device != self.device or not z.is_contiguous()
                or z.numel() != self.nchol * o * o):
            raise ValueError("z must be a contiguous float64 device tensor of nchol*nocc*nocc elements")
        return o

    def exchange(self, corbs, precision="fp64", z=None) -> torch.Tensor:
        """K_d = sum_J (L^J C_d)(L^J C_d)^T in LT storage for each orbital
        matrix C_d (nbas, o_d) on the device (C-order = Porb memory).  With z
        (one orbital matrix only; (nchol, o, o) float64 on the device) the
        same pass also writes Z^J = C^T L^J C for grad_vectors(z=...)."""
        cs = [_dev_f64(c, self.device) for c in corbs]
        if z is not None:
            o = self._z_out(cs, z)
            k = torch.empty((1, self.npair), dtype=torch.float64, device=self.device)
            self._check(self.lib.mjk_exchange_z(self._h, o, _ptr(cs[0]), _ptr(k),
                                                precision_code(precision), _ptr(z)))
            return k
        nk = len(cs)
        nocc = np.array([c.shape[1] if c.dim() == 2 else 0 for c in cs], dtype=np.int32)
        ptrs = (ctypes.c_void_p * nk)(*[c.data_ptr() for c in cs])
        k = torch.empty((nk, self.npair), dtype=torch.float64, device=self.device)
        self._check(self.lib.mjk_exchange(self._h, nk, _ptr(nocc), ptrs, _ptr(k), _ld(k),
                                          precision_code(precision)))
        return k

    def coulomb_exchange(self, dlt: torch.Tensor, corbs, want_v: bool = False, z=None):
        """J of one folded LT density dlt (npair,) and the FP64 K_d of each
        orbital matrix C_d in one pass over the vectors (the V = L^T D pass
        runs inside the half-transform).  Returns (J (1, npair), K (nk, npair))
        and V (nchol, 1) when want_v.  With z (one orbital matrix only) the
        pass also writes Z^J = C^T L^J C, as exchange(z=...)."""
        d = _dev_f64(dlt, self.device).reshape(1, -1)
        if d.shape[1] != self.npair:
            raise ValueError("dlt must hold one folded density of nbas(nbas+1)/2 elements")
        cs = [_dev_f64(c, self.device) for c in corbs]
        if z is not None:
            o = self._z_out(cs, z)
            j = torch.empty((1, self.npair), dtype=torch.float64, device=self.device)
            k = torch.empty((1, self.npair), dtype=torch.float64, device=self.device)
            v = torch.empty((self.nchol, 1), dtype=torch.float64, device=self.device) if want_v else None
            self._check(self.lib.mjk_coulomb_exchange_z(self._h, _ptr(d), _ptr(j), o, _ptr(cs[0]),
                                                        _ptr(k), _ptr(v), _ptr(z)))
            return (j, k, v) if want_v else (j, k)
        nk = len(cs)
        nocc = np.array([c.shape[1] if c.dim() == 2 else 0 for c in cs], dtype=np.int32)
        ptrs = (ctypes.c_void_p * nk)(*[c.data_ptr() for c in cs])
        j = torch.empty((1, self.npair), dtype=torch.float64, device=self.device)
        k = torch.empty((nk, self.npair), dtype=torch.float64, device=self.device)
        v = torch.empty((self.nchol, 1), dtype=torch.float64, device=self.device) if want_v else None
        self._check(self.lib.mjk_coulomb_exchange(self._h, _ptr(d), _ld(d), _ptr(j), _ld(j), nk,
                                                  _ptr(nocc), ptrs, _ptr(k), _ld(k), _ptr(v)))
        return (j, k, v) if want_v else (j, k)

    def exchange_mixed(self, corbs, j_fp32: int, j_bf16: int) -> torch.Tensor:
        """Exchange with vectors [0, j_fp32) in FP64, [j_fp32, j_bf16) in FP32
        and [j_bf16, nchol) in split-BF16, reduced in FP64."""
        cs = [_dev_f64(c, self.device) for c in corbs]
        nk = len(cs)
        nocc = np.array([c.shape[1] if c.dim() == 2 else 0 for c in cs], dtype=np.int32)
        ptrs = (ctypes.c_void_p * nk)(*[c.data_ptr() for c in cs])
        k = torch.empty((nk, self.npair), dtype=torch.float64, device=self.device)
        self._check(self.lib.mjk_exchange_mixed(self._h, nk, _ptr(nocc), ptrs, _ptr(k), _ld(k),
                                                int(j_fp32), int(j_bf16)))
        return k

    def cd_incore(self, x: torch.Tensor, thr: float, mxvec=None):
        """CD_InCore on the device: returns (irc, V) with V (n, numcho) and
        x @ x.T-style reconstruction V V^T ~ x to within thr on the diagonal.
        x (n, n) symmetric (a copy is decomposed)."""
        n = x.shape[0]
        xc = _dev_f64(x, self.device).t().contiguous()  # column-major copy (symmetric anyway)
        mx = n if mxvec is None else int(mxvec)
        v = torch.zeros((max(mx, 1), n), dtype=torch.float64, device=self.device)  # column-major (n, mx)
        num = ctypes.c_int32(0)
        irc = ctypes.c_int32(0)
        self._check(self.lib.mjk_cd_incore(self._h, int(n), _ptr(xc), _ptr(v), mx, ctypes.byref(num),
                                           float(thr), ctypes.byref(irc)))
        return int(irc.value), v[:num.value].t()

    def exchange_screened(self, corbs, tau: float, thr_x: float = 0.0, precision="fp64") -> torch.Tensor:
        """Exchange with tile screening (see mjk_exchange_screened)."""
        cs = [_dev_f64(c, self.device) for c in corbs]
        nk = len(cs)
        nocc = np.array([c.shape[1] if c.dim() == 2 else 0 for c in cs], dtype=np.int32)
        ptrs = (ctypes.c_void_p * nk)(*[c.data_ptr() for c in cs])
        k = torch.empty((nk, self.npair), dtype=torch.float64, device=self.device)
        self._check(self.lib.mjk_exchange_screened(self._h, nk, _ptr(nocc), ptrs, _ptr(k), _ld(k),
                                                   precision_code(precision), float(tau),
                                                   float(thr_x)))
        return k

    def screening_stats(self) -> dict:
        """Work items of the last exchange: K tiles x slices and half-transform tiles."""
        v = [ctypes.c_int64(0) for _ in range(4)]
        self._check(self.lib.mjk_screening_stats(self._h, *[ctypes.byref(x) for x in v]))
        return {"k_items": v[0].value, "k_skipped": v[1].value, "x_tiles": v[2].value,
                "x_skipped": v[3].value}

    def set_lk_screening(self, enable: bool, thrcom: float = 1e-6) -> None:
        """LK screening of the LK drop-ins (the reference's); thrcom is the
        decomposition threshold ThrCom."""
        self._check(self.lib.mjk_set_lk_screening(self._h, int(bool(enable)), float(thrcom)))

    def set_shells(self, nbassh) -> None:
        """Shell sizes nBasSh (consecutive AOs per shell, in order)."""
        sh = np.ascontiguousarray(nbassh, dtype=np.int32)
        self._check(self.lib.mjk_set_shells(self._h, len(sh), _ptr(sh)))

    def set_lk_diag(self, diag, maxred: int = 0, estimate: bool = False, update: bool = True) -> None:
        """The integral diagonal (LT order) each LK call starts from, MaxRed and
        the Estimate / Update switches."""
        d = _host_f64(diag).reshape(-1)
        if d.size != self.npair:
            raise ValueError("diag must hold nbas(nbas+1)/2 values")
        self._check(self.lib.mjk_set_lk_diag(self._h, _ptr(d), int(maxred), int(bool(estimate)),
                                             int(bool(update))))

    def lk_stats(self) -> np.ndarray:
        """[listed shells, half-transformed shell pairs, K blocks, batches] of
        the last screened LK call."""
        st = np.zeros(4, dtype=np.int64)
        self._check(self.lib.mjk_lk_stats(self._h, _ptr(st)))
        return st

    def vector_norms(self) -> torch.Tensor:
        """(nchol,) squared Frobenius norms of the symmetric vectors L^J."""
        w = torch.empty(self.nchol, dtype=torch.float64, device=self.device)
        self._check(self.lib.mjk_vector_norms(self._h, _ptr(w)))
        return w

    def fock_update(self, flt: torch.Tensor, jlt=None, jsrc=None, fj=None, klt=None, fk=None):
        """flt[d] += fj[d] * jlt[jsrc[d]] + fk[d] * klt[d] (in place)."""
        nd = flt.shape[0]
        js = np.ascontiguousarray(jsrc if jsrc is not None else [-1] * nd, dtype=np.int32)
        a = np.ascontiguousarray(fj if fj is not None else [0.0] * nd, dtype=np.float64)
        b = np.ascontiguousarray(fk if fk is not None else [0.0] * nd, dtype=np.float64)
        self._check(self.lib.mjk_fock_update(
            self._h, nd, _ptr(flt), _ld(flt), _ptr(jlt), _ld(jlt) if jlt is not None else 0,
            _ptr(js), _ptr(a), _ptr(klt), _ld(klt) if klt is not None else 0, _ptr(b)))
        return flt

    def energy(self, h, dlt, g) -> torch.Tensor:
        """sum_LT h*dlt + 1/2 sum_LT g*dlt (device scalar)."""
        e = torch.empty(1, dtype=torch.float64, device=self.device)
        self._check(self.lib.mjk_energy(self._h, _ptr(h), _ptr(dlt), _ptr(g), _ptr(e)))
        return e

    def puvx(self, porb3: torch.Tensor, cmo: torch.Tensor, nfro: int, norb: int) -> torch.Tensor:
        """(pw|xy) for active orbitals porb3 (nbas, naorb); cmo (nbas, nbas),
        cmo[a, p] = CMO(a,p).  Returns W_PWXY as a flat tensor."""
        c3 = _dev_f64(porb3, self.device)
        cm = _dev_f64(cmo, self.device).t().contiguous()
        na = c3.shape[1]
        w = torch.empty(norb * na * (na * (na + 1) // 2), dtype=torch.float64, device=self.device)
        self._check(self.lib.mjk_puvx(self._h, na, _ptr(c3), _ptr(cm), int(nfro), int(norb), _ptr(w)))
        return w

    def density(self, corb: torch.Tensor, scale: float = 1.0) -> torch.Tensor:
        """Folded LT of scale * C C^T for C (nbas, nocc) on the device."""
        c = _dev_f64(corb, self.device)
        d = torch.empty((1, self.npair), dtype=torch.float64, device=self.device)
        self._check(self.lib.mjk_density(self._h, c.shape[1], _ptr(c), float(scale), _ptr(d)))
        return d

    def grad_orbitals(self, corb: torch.Tensor, flt: torch.Tensor, scale: float) -> torch.Tensor:
        """scale * F C with F in plain LT storage."""
        c = _dev_f64(corb, self.device)
        dc = torch.empty_like(c)
        self._check(self.lib.mjk_grad_orbitals(self._h, c.shape[1], _ptr(c), _ptr(flt), float(scale), _ptr(dc)))
        return dc

    def grad_vectors(self, corb: torch.Tensor, v: torch.Tensor, alpha: float, beta: float,
                     p_lt=None, out=None, z=None, fp32_from=None) -> torch.Tensor:
        """(nchol, npair) packed gradient s_ab (alpha V_J P_ab + beta (C Z^J C^T)_ab).
        z: the Z^J = C^T L^J C an exchange(z=...) or coulomb_exchange(z=...)
        of the same orbitals wrote (no half-transform then); None recomputes.
        fp32_from (with z): vectors J >= fp32_from form C Z^J C^T in FP32."""
        c = _dev_f64(corb, self.device)
        vv = v.to(torch.float64).contiguous()
        dl = out if out is not None else torch.empty((self.nchol, self.npair), dtype=torch.float64,
                                                     device=self.device)
        if z is not None:
            o = self._z_out([c], z)
            if fp32_from is not None and int(fp32_from) < self.nchol:
                self._check(self.lib.mjk_grad_vectors_z_mixed(
                    self._h, o, _ptr(c), _ptr(p_lt), _ptr(vv), _ptr(z), float(alpha), float(beta),
                    _ptr(dl), _ld(dl), int(fp32_from)))
                return dl
            self._check(self.lib.mjk_grad_vectors_z(self._h, o, _ptr(c), _ptr(p_lt), _ptr(vv), _ptr(z),
                                                    float(alpha), float(beta), _ptr(dl), _ld(dl)))
            return dl
        if fp32_from is not None:
            raise ValueError("grad_vectors: fp32_from needs the saved z")
        self._check(self.lib.mjk_grad_vectors(self._h, c.shape[1], _ptr(c), _ptr(p_lt), _ptr(vv),
                                              float(alpha), float(beta), _ptr(dl), _ld(dl)))
        return dl

    def device_vectors(self) -> torch.Tensor:
        """(nchol, npair) float64 view of the resident vector store (zero copy);
        writing it replaces the vectors."""
        p = ctypes.c_void_p()
        ld = ctypes.c_int64()
        self._check(self.lib.mjk_device_vectors(self._h, ctypes.byref(p), ctypes.byref(ld)))

        class _View:
            __cuda_array_interface__ = {
                "shape": (self.nchol, self.npair), "strides": (8 * ld.value, 8),
                "typestr": "<f8", "data": (p.value, False), "version": 3}

        t = torch.as_tensor(_View(), device=self.device)
        self._resident = ("device", t.data_ptr())
        return t

    def set_tuning(self, tile_syrk: int = 0, split_target: int = 0, ht_ta: int = 0) -> None:
        self._check(self.lib.mjk_set_tuning(self._h, int(tile_syrk), int(split_target), int(ht_ta)))

    def autotune(self, corbs, precision="fp64") -> float:
        """Times the exchange tilings for these orbital sets; keeps the fastest."""
        cs = [_dev_f64(c, self.device) for c in corbs]
        nk = len(cs)
        nocc = np.array([c.shape[1] for c in cs], dtype=np.int32)
        ptrs = (ctypes.c_void_p * nk)(*[c.data_ptr() for c in cs])
        k = torch.empty((nk, self.npair), dtype=torch.float64, device=self.device)
        best = ctypes.c_double()
        self._check(self.lib.mjk_autotune(self._h, nk, _ptr(nocc), ptrs, _ptr(k), _ld(k),
                                          precision_code(precision), ctypes.byref(best)))
        return best.value

    def ensure_vectors(self, L: torch.Tensor) -> None:
        """Uploads L (nchol, npair) unless this exact tensor version is resident."""
        res = getattr(self, "_resident", None)
        if res is not None and res[0] == "device" and res[1] == L.data_ptr():
            return  # L is the context's own store
        key = (L.data_ptr(), L._version, tuple(L.shape))
        if res != key:
            self.upload_packed(L.detach())
            self._resident = key

    # ------------------------------------------------------------ measurement
    def set_profiling(self, on: bool):
        self._check(self.lib.mjk_set_profiling(self._h, int(bool(on))))

    def measure_peak(self, kind: int) -> float:
        """Measured device peak: 0 FP64 MFMA, 1 FP32 MFMA, 2 BF16 MFMA, 3 FP64 FMA (TFLOP/s), 4 copy (TB/s)."""
        v = ctypes.c_double()
        self._check(self.lib.mjk_measure_peak(self._h, int(kind), ctypes.byref(v)))
        return v.value

    def phase_times(self):
        ms = (ctypes.c_double * len(_lib.PHASES))()
        calls = (ctypes.c_int64 * len(_lib.PHASES))()
        self._check(self.lib.mjk_phase_times(self._h, ms, calls))
        return {p: (ms[i], calls[i]) for i, p in enumerate(_lib.PHASES)}


class CholeskyFiles:
    """The reference's Cholesky index files read on the host (Cho_X_Init,
    nSym = 1..8): CHORST (restart info, InfVec) and CHRED (reduced sets), the
    element maps rebuilt from isoshl (the runfile's ISOSHL, 1-based shell of
    each SO, SOs irrep by irrep) into the C1 triangle of all nBasT SOs.  No
    device is needed to read; ``load`` feeds a JKContext (one CHVEC path, or
    one per irrep)."""

    def __init__(self, chorst: str, chred: str, isoshl):
        self.lib = _lib.load()
        sh = np.ascontiguousarray(isoshl, dtype=np.int32)
        self._keep = sh
        h = ctypes.c_void_p()
        rc = self.lib.mjk_choinfo_read(ctypes.byref(h), os.fsencode(chorst), os.fsencode(chred), _ptr(sh))
        self._h = h
        if rc != 0:
            msg = self.lib.mjk_choinfo_error(h).decode()
            self.close()
            raise MjkError(rc, msg)
        nb, mr, nsh, av = (ctypes.c_int32() for _ in range(4))
        nc, thr = ctypes.c_int64(), ctypes.c_double()
        self.lib.mjk_choinfo_dims(h, ctypes.byref(nb), ctypes.byref(nc), ctypes.byref(mr), ctypes.byref(nsh),
                                  ctypes.byref(av), ctypes.byref(thr))
        self.nbas, self.numcho, self.maxred = nb.value, nc.value, mr.value
        self.nshell, self.adrvec, self.thrcom = nsh.value, av.value, thr.value
        ns = ctypes.c_int32()
        nbs = np.zeros(8, dtype=np.int32)
        ncs = np.zeros(8, dtype=np.int64)
        self.lib.mjk_choinfo_symmetry(h, ctypes.byref(ns), _ptr(nbs), _ptr(ncs))
        self.nsym = ns.value
        self.nbas_sym = [int(v) for v in nbs[:self.nsym]]
        self.numcho_sym = [int(v) for v in ncs[:self.nsym]]

    def reduced_set(self, jred: int, isym: int = 1) -> np.ndarray:
        """0-based LT index of each element of reduced set jred (0-based) in
        irrep isym (1-based)."""
        nrs = ctypes.c_int64()
        k = (int(isym) - 1) * self.maxred + int(jred)
        if not 1 <= int(isym) <= self.nsym or \
                self.lib.mjk_choinfo_reduced_set(self._h, k, ctypes.byref(nrs), None) != 0:
            raise MjkError(1, f"no reduced set {jred} in irrep {isym}")
        m = np.zeros(nrs.value, dtype=np.int64)
        self.lib.mjk_choinfo_reduced_set(self._h, k, ctypes.byref(nrs), _ptr(m))
        return m

    def infvec(self, col: int) -> np.ndarray:
        v = np.zeros(self.numcho, dtype=np.int64)
        if self.numcho and self.lib.mjk_choinfo_infvec(self._h, int(col), _ptr(v)) != 0:
            raise MjkError(1, f"bad InfVec column {col}")
        return v

    def load(self, jk: "JKContext", chvec, chodiag: str = None, split_bytes: int = 0) -> None:
        """chvec: the CHVEC file (nSym = 1) or a list of one per irrep."""
        diag = os.fsencode(chodiag) if chodiag else None
        if isinstance(chvec, (list, tuple)):
            names = [os.fsencode(p) for p in chvec]
            arr = (ctypes.c_char_p * len(names))(*names)
            jk._check(self.lib.mjk_load_choinfo_sym(jk._h, self._h, arr, len(names), diag, int(split_bytes)))
            return
        jk._check(self.lib.mjk_load_choinfo(jk._h, self._h, os.fsencode(chvec), diag, int(split_bytes)))

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self.lib.mjk_choinfo_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
