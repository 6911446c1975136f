"""The callers' side of the J/K build: the dispatch of ChoSCF_Drv and CHO_CAS_DRV
on top of the drop-ins, with the density decomposition (DECO) on the device.

* ``choscf_fock`` follows src/scf/choscf_drv.F90 (nSym = 1, ALGO 3 -> CHO_FSCF,
  ALGO 4 -> CHO_LK_SCF): RHF (nD = 1, :72-246) and UHF (nD = 2, :250-440).
  DECO decomposes each square density with CD_InCore at 1e-8 * max diagonal
  (:111-115, :313-320) and halves the RHF exchange factor because the vectors
  reproduce the doubly occupied density (:101,135); without DECO the occupied
  MO columns are used with the full ExFac.  UHF sends the total density to
  Coulomb and the alpha/beta densities to exchange with FactX = ExFac
  (:276-298).
* ``cho_cas_drv`` follows src/fock_util/cho_cas_drv.F90:14-330: folds DI and
  DA (:104-105), FactXI = -1, or -1/2 with DECO of DI at Thr = 1e-12
  (:21,107-129); the active density is always decomposed (:202-227); then
  CHO_FMCSCF (ALGO 1) or CHO_LK_CASSCF (LK) on zeroed FI/FA (:268-286).

All arrays are host numpy arrays (the reference's calling convention);
densities are square (n, n), CMO is (n, n) with CMO[a, p] = CMO(a,p).
"""
from __future__ import annotations

import numpy as np
import torch

from .fock import JKContext


def _fold(dsq: np.ndarray) -> np.ndarray:
    """Square symmetric -> LT with off-diagonal elements doubled
    (src/misc_util/fold.F90:14; a(a+1)/2 + b order)."""
    n = dsq.shape[0]
    i, j = np.tril_indices(n)
    out = dsq[i, j] + dsq[j, i]
    out[i == j] *= 0.5
    return np.ascontiguousarray(out)


def _deco(jk: JKContext, dsq: np.ndarray, thr: float) -> np.ndarray:
    """CD_InCore of a square density on the device; returns V (n, numv)."""
    irc, v = jk.cd_incore(torch.from_numpy(np.ascontiguousarray(dsq, dtype=np.float64)), thr)
    if irc != 0:
        raise RuntimeError(f"CD_InCore rc={irc}: ill-defined density decomposition")
    return v.cpu().numpy()


def choscf_fock(jk: JKContext, dsq, exfac: float = 1.0, algo: int = 3, deco: bool = True,
                cmo=None, nocc=None, dsq_beta=None, cmo_beta=None, nocc_beta=None,
                flt=None, nbassh=None, precision="fp64") -> np.ndarray:
    """Two-electron Fock matrices in LT storage, (1, npair) for RHF or (2, npair)
    (alpha, beta) for UHF when ``dsq_beta`` is given; added to ``flt`` if given.

    RHF: dsq is the total (doubly occupied) density.  UHF: dsq / dsq_beta are
    the alpha / beta densities.  Without DECO, cmo/nocc (and cmo_beta/nocc_beta)
    give the occupied MO columns."""
    if algo not in (3, 4):
        raise ValueError("only ALGO 3 (CHO_FSCF) and 4 (CHO_LK_SCF) run on the device")
    uhf = dsq_beta is not None
    n = jk.nbas
    nden = 2 if uhf else 1
    out = np.zeros((nden, jk.npair)) if flt is None else np.array(flt, dtype=np.float64).reshape(nden, -1)
    out = np.ascontiguousarray(out)
    dens = [np.asarray(dsq, dtype=np.float64)] + ([np.asarray(dsq_beta, dtype=np.float64)] if uhf else [])
    dlt = _fold(dens[0] + dens[1]) if uhf else _fold(dens[0])
    if exfac == 0.0:  # pure DFT: CHO_FOCK_DFT_RED writes J into FLT(1) (:95, :302)
        f1 = np.ascontiguousarray(out[0])
        jk.cho_fock_dft_red(dlt, f1)
        out[0] = f1
        return out
    if deco:
        orbs = [_deco(jk, d, 1e-8 * float(np.max(np.diag(d)))) for d in dens]
        xfac = exfac if uhf else 0.5 * exfac
    else:
        if uhf:
            orbs = [np.asarray(cmo)[:, :nocc], np.asarray(cmo_beta)[:, :nocc_beta]]
        else:
            orbs = [np.asarray(cmo)[:, :nocc]]
        xfac = exfac
    nocc_d = [o.shape[1] for o in orbs]
    if algo == 3:
        jk.cho_fscf(nden, out, [0] * nden, nocc_d, [np.ascontiguousarray(o) for o in orbs], dlt, xfac,
                    precision)
    else:
        klt = np.zeros_like(out)
        jk.cho_lk_scf(nden, out, klt, [0] * nden, nocc_d, orbs, dlt, xfac, nbassh=nbassh,
                      precision=precision)
    return out


def cho_cas_drv(jk: JKContext, cmo, di_sq, da_sq, nfro: int, nish: int, nash: int,
                deco: bool = True, do_active: bool = True, exfac: float = 1.0, lk: bool = False,
                precision="fp64"):
    """Returns (FI, FA, W_PWXY): FI/FA in LT storage (npair,), W_PWXY for the
    active orbitals nfro+nish .. nfro+nish+nash-1 over all norb = nmo - nfro
    MOs after the frozen ones."""
    n = jk.nbas
    cmo = np.asarray(cmo, dtype=np.float64)
    if cmo.shape != (n, n):
        raise ValueError("cmo must be the square (nbas, nbas) MO coefficient matrix")
    dlt = np.stack([_fold(np.asarray(di_sq, dtype=np.float64)), _fold(np.asarray(da_sq, dtype=np.float64))])
    thr = 1.0e-12
    nocc = nfro + nish
    if deco:
        factxi = -0.5
        porb1 = _deco(jk, di_sq, thr) if nocc > 0 else np.zeros((n, 0))
    else:
        factxi = -1.0
        porb1 = cmo[:, :nocc]
    porb3 = cmo[:, nocc:nocc + nash]
    if do_active and nash > 0:
        porb2 = _deco(jk, da_sq, thr)
    else:
        porb2 = np.zeros((n, 0))
    flt = np.zeros((2, jk.npair))
    norb = cmo.shape[1] - nfro
    nxy = nash * (nash + 1) // 2
    w = np.zeros(max(1, norb * nash * nxy))
    if lk:  # DoLocK: CHO_LK_CASSCF with MSQ and CVa (cho_cas_drv.F90:281-286)
        jk.cho_lk_casscf(dlt, flt, porb1, w, factxi, porb1.shape[1], nash, porb2.shape[1], porb3,
                         porb2, do_active, cmo, nfro, norb, exfac=exfac, precision=precision)
    else:   # ALGO 1: CHO_FMCSCF (cho_cas_drv.F90:273-277)
        jk.cho_fmcscf(flt, 0, porb1.shape[1], nash, factxi, dlt, do_active,
                      np.ascontiguousarray(porb1), porb2.shape[1], np.ascontiguousarray(porb2),
                      np.ascontiguousarray(porb3), w, cmo, nfro, norb, exfac, precision)
    return flt[0], flt[1], w[:norb * nash * nxy]
