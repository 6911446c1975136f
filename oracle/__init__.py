"""CPU FP64 oracle for the Cholesky J/K Fock build.

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by ``__graft_entry__.smoke()``
as the parity reference, and by ``bench.py``'s ``cpu_baseline`` leg -- never by the
product package.  The C restatement lives in ``mjk_oracle.c`` (each function
cites the OpenMolcas routine it follows); this module is its ctypes binding plus
a few numpy helpers that the checks share.

Parity status of each function (how it is pinned, see tests/test_oracle.py):
  fold/unfold/itri      -- textbook identities and the reference's formulas
  swap_full2rs/rs2full  -- round trip on random nested reduced sets
  cho_vtra2/vtra0       -- dense brute force X = C^T L (numpy einsum)
  dgemm_tri_tn/nt       -- dense numpy products, lower triangle
  cho_fscf              -- dense brute force J/K, batching invariance,
                           reduced-set invariance, 4FC gradient identity
  cho_fmcscf            -- dense brute force FI/FA/(pw|xy)
  cho_fock_dft_red      -- dense brute force J over several reduced sets
                           (FLT overwritten)
  cho_lk_scf            -- dense brute force J[P] - FactXI K, shell-pair KLT
                           layout against an index map built from the block
                           rules, equality with cho_fscf (unscreened)
  cho_lk_casscf         -- dense brute force FI/FA with ExFac-scaled factors,
                           (pw|xy), equality with cho_fmcscf at ExFac = 1
  cho_lk_scf_screened / cho_lk_casscf_screened (mjk_oracle_lk.c)
                        -- dmpk = 0 equals the unscreened routines; the
                           kept shell lists, half-transformed shell pairs
                           and K blocks equal an independent per-batch
                           evaluation of the reference's inequalities
                           (tests/_lk_brute.py); skipped blocks are
                           bounded by the thresholds on localized inputs
  cd_incore             -- ||D - V V^T||_max <= thr, exact rank, numpy cholesky
  energy                -- closed form tr(hP) + 1/2 tr(PJ) - 1/4 tr(PK)
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libmjk_oracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_i64p = ctypes.POINTER(ctypes.c_int64)
_i32p = ctypes.POINTER(ctypes.c_int32)


class _LK(ctypes.Structure):
    _fields_ = [
        ("nshell", ctypes.c_int32),
        ("nbassh", _i32p),
        ("diag", _dp),
        ("thrcom", ctypes.c_double),
        ("maxred", ctypes.c_int32),
        ("estimate", ctypes.c_int32),
        ("update", ctypes.c_int32),
        ("nscreen", ctypes.c_int32),
    ]


class _Chol(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int32),
        ("nred", ctypes.c_int32),
        ("nrs", _i64p),
        ("rs_off", _i64p),
        ("rs2lt", _i64p),
        ("nvec", ctypes.c_int64),
        ("vred", _i32p),
        ("voff", _i64p),
        ("lvec", _dp),
    ]


def build() -> str:
    """Compiles the oracle with its Makefile (gcc) and returns the library path."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.ora_set_threads.argtypes = [ctypes.c_int32]
        L.ora_set_dgemv.argtypes = [ctypes.c_void_p]
        L.ora_get_threads.restype = ctypes.c_int32
        L.ora_itri.restype = ctypes.c_int64
        L.ora_itri.argtypes = [ctypes.c_int64, ctypes.c_int64]
        L.ora_fold.argtypes = [ctypes.c_int32, _dp, _dp]
        L.ora_unfold.argtypes = [ctypes.c_int32, _dp, _dp]
        L.ora_swap_full2rs.argtypes = [ctypes.c_int64, _i64p, ctypes.c_int32, ctypes.c_int64, _dp, _dp]
        L.ora_swap_rs2full.argtypes = [ctypes.c_int64, _i64p, ctypes.c_int32, ctypes.c_int64, _dp, _dp,
                                       ctypes.c_int32]
        L.ora_dgemm_tri_tn.argtypes = [ctypes.c_int64, ctypes.c_int64, ctypes.c_double, _dp, ctypes.c_int64,
                                       _dp, ctypes.c_int64, ctypes.c_double, _dp]
        L.ora_dgemm_tri_nt.argtypes = L.ora_dgemm_tri_tn.argtypes
        L.ora_cho_vtra2.argtypes = [ctypes.c_int32, ctypes.c_int64, _i64p, ctypes.c_int64, _dp,
                                    ctypes.c_int32, _dp, _dp]
        L.ora_cho_vtra0.argtypes = L.ora_cho_vtra2.argtypes
        L.ora_cho_fscf.restype = ctypes.c_int32
        L.ora_cho_fscf.argtypes = [ctypes.POINTER(_Chol), ctypes.c_int32, _dp, _i32p, _i32p,
                                   ctypes.POINTER(_dp), _dp, ctypes.c_double, ctypes.c_int64]
        L.ora_cho_fock_dft_red.restype = ctypes.c_int32
        L.ora_cho_fock_dft_red.argtypes = [ctypes.POINTER(_Chol), _dp, _dp, ctypes.c_int64]
        L.ora_cho_lk_scf.restype = ctypes.c_int32
        L.ora_cho_lk_scf.argtypes = [ctypes.POINTER(_Chol), ctypes.c_int32, _dp, _dp, _i32p, _i32p,
                                     ctypes.POINTER(_dp), _dp, ctypes.c_double, ctypes.c_int32,
                                     _i32p, ctypes.c_int64]
        L.ora_cho_fmcscf.restype = ctypes.c_int32
        L.ora_cho_fmcscf.argtypes = [ctypes.POINTER(_Chol), _dp, ctypes.c_int32, ctypes.c_int32,
                                     ctypes.c_int32, ctypes.c_double, _dp, ctypes.c_int32, _dp,
                                     ctypes.c_int32, _dp, _dp, _dp, _dp, ctypes.c_int32,
                                     ctypes.c_int32, ctypes.c_double, ctypes.c_int64]
        L.ora_cho_lk_casscf.restype = ctypes.c_int32
        L.ora_cho_lk_casscf.argtypes = [ctypes.POINTER(_Chol), _dp, _dp, _dp, _dp, ctypes.c_double,
                                        ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _dp, _dp,
                                        ctypes.c_int32, _dp, ctypes.c_int32, ctypes.c_int32,
                                        ctypes.c_double, ctypes.c_int64]
        L.ora_lk_thresholds.restype = None
        L.ora_lk_thresholds.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int32,
                                        _i32p, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, _dp, _dp]
        L.ora_cho_lk_scf_scr.restype = ctypes.c_int32
        L.ora_cho_lk_scf_scr.argtypes = [ctypes.POINTER(_Chol), ctypes.c_int32, _dp, _dp, _i32p, _i32p,
                                         ctypes.POINTER(_dp), _dp, ctypes.c_double, ctypes.c_int32,
                                         ctypes.c_double, ctypes.c_double, ctypes.POINTER(_LK),
                                         ctypes.c_int64, _i64p]
        L.ora_cho_lk_casscf_scr.restype = ctypes.c_int32
        L.ora_cho_lk_casscf_scr.argtypes = [ctypes.POINTER(_Chol), _dp, _dp, _dp, _dp, ctypes.c_double,
                                            ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _dp, _dp,
                                            ctypes.c_int32, ctypes.c_int32, ctypes.c_double,
                                            ctypes.c_double, _dp, ctypes.c_int32, ctypes.c_int32,
                                            ctypes.c_double, ctypes.POINTER(_LK), ctypes.c_int64, _i64p]
        L.ora_cd_incore.restype = ctypes.c_int32
        L.ora_cd_incore.argtypes = [_dp, ctypes.c_int32, _dp, ctypes.c_int32, _i32p, ctypes.c_double]
        L.ora_energy.restype = ctypes.c_double
        L.ora_energy.argtypes = [ctypes.c_int64, _dp, _dp, _dp, ctypes.c_double]
        _lib = L
    return _lib


def set_threads(n: int) -> None:
    """OpenMP threads of the restatement (1: a single core)."""
    lib().ora_set_threads(int(n))


def get_threads() -> int:
    return int(lib().ora_get_threads())


def use_blas(on: bool = True) -> str:
    """Timing only: route the DGEMV calls of the restatement (the J pair and
    DGEMM_Tri's per-column DGEMVs) through scipy's OpenBLAS dgemv, as the
    reference calls its linked BLAS; its threads follow threadpoolctl.  Off
    (the default) keeps the plain loops every check uses.  Returns the BLAS
    build in use, or "" when off."""
    if not on:
        lib().ora_set_dgemv(None)
        return ""
    from scipy.linalg import cython_blas

    cap = cython_blas.__pyx_capi__["dgemv"]
    get_name = ctypes.pythonapi.PyCapsule_GetName
    get_name.restype = ctypes.c_char_p
    get_name.argtypes = [ctypes.py_object]
    get_ptr = ctypes.pythonapi.PyCapsule_GetPointer
    get_ptr.restype = ctypes.c_void_p
    get_ptr.argtypes = [ctypes.py_object, ctypes.c_char_p]
    lib().ora_set_dgemv(get_ptr(cap, get_name(cap)))
    try:
        from threadpoolctl import threadpool_info

        info = [i for i in threadpool_info() if i.get("user_api") == "blas"]
        if info:
            return f"{info[0].get('internal_api', 'blas')} {info[0].get('version', '')} (scipy)".strip()
    except ImportError:
        pass
    return "scipy BLAS"


def _p(a: np.ndarray, ct=_dp):
    return a.ctypes.data_as(ct)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


# ---------------------------------------------------------------- vector store
class CholStore:
    """Reduced sets + vectors ordered by reduced set (the state Cho_X_Init
    leaves behind).  ``sets`` is a list of 0-based LT index arrays; ``vectors``
    a list of (set_id, values) pairs, set ids nondecreasing."""

    def __init__(self, n: int, sets, vectors):
        self.n = int(n)
        self.sets = [np.ascontiguousarray(s, dtype=np.int64) for s in sets]
        self.nrs = np.array([len(s) for s in self.sets], dtype=np.int64)
        self.rs_off = np.concatenate([[0], np.cumsum(self.nrs)[:-1]]).astype(np.int64)
        self.rs2lt = np.concatenate(self.sets).astype(np.int64) if self.sets else np.zeros(0, np.int64)
        self.vred = np.array([r for r, _ in vectors], dtype=np.int32)
        if len(self.vred) and np.any(np.diff(self.vred) < 0):
            raise ValueError("vectors must be ordered by reduced set")
        vals = [np.ascontiguousarray(v, dtype=np.float64) for _, v in vectors]
        for r, v in zip(self.vred, vals):
            if v.shape != (self.nrs[r],):
                raise ValueError("vector length does not match its reduced set")
        lens = np.array([len(v) for v in vals], dtype=np.int64)
        self.voff = np.concatenate([[0], np.cumsum(lens)[:-1]]).astype(np.int64)
        self.lvec = np.concatenate(vals) if vals else np.zeros(1)
        self._c = _Chol(self.n, len(self.sets), _p(self.nrs, _i64p), _p(self.rs_off, _i64p),
                        _p(self.rs2lt, _i64p), len(vals), _p(self.vred, _i32p),
                        _p(self.voff, _i64p), _p(self.lvec))

    @classmethod
    def full(cls, L_packed: np.ndarray) -> "CholStore":
        """One reduced set covering every LT pair; L_packed is (M, n(n+1)/2)."""
        M, npair = L_packed.shape
        n = int(round((np.sqrt(8 * npair + 1) - 1) / 2))
        idx = np.arange(npair, dtype=np.int64)
        return cls(n, [idx], [(0, L_packed[j]) for j in range(M)])

    def packed(self) -> np.ndarray:
        """Vectors expanded to full LT storage, (M, n(n+1)/2)."""
        npair = self.n * (self.n + 1) // 2
        out = np.zeros((len(self.vred), npair))
        for j, r in enumerate(self.vred):
            off = self.voff[j]
            out[j, self.sets[r]] = self.lvec[off:off + self.nrs[r]]
        return out


# ---------------------------------------------------------------- wrappers
def itri(i: int, j: int) -> int:
    return int(lib().ora_itri(i, j))


def fold(a_sq: np.ndarray) -> np.ndarray:
    a = _f64(a_sq)
    n = a.shape[0]
    # the C routine reads A(i,j) column-major; a symmetric input makes the
    # distinction moot, a general one is passed transposed
    at = _f64(a.T)
    out = np.zeros(n * (n + 1) // 2)
    lib().ora_fold(n, _p(at), _p(out))
    return out


def unfold(a_lt: np.ndarray, n: int) -> np.ndarray:
    a = _f64(a_lt)
    out = np.zeros((n, n))
    lib().ora_unfold(n, _p(a), _p(out))
    return out


def swap_full2rs(rs2lt: np.ndarray, xlt: np.ndarray) -> np.ndarray:
    m = np.ascontiguousarray(rs2lt, dtype=np.int64)
    x = _f64(xlt)
    out = np.zeros(len(m))
    lib().ora_swap_full2rs(len(m), _p(m, _i64p), 1, x.shape[-1], _p(x), _p(out))
    return out


def swap_rs2full(rs2lt: np.ndarray, xrs: np.ndarray, xlt: np.ndarray, add: bool) -> np.ndarray:
    m = np.ascontiguousarray(rs2lt, dtype=np.int64)
    x = _f64(xrs)
    out = _f64(xlt).copy()
    lib().ora_swap_rs2full(len(m), _p(m, _i64p), 1, out.shape[-1], _p(x), _p(out), int(add))
    return out


def cho_vtra2(n: int, rs2lt: np.ndarray, lrs: np.ndarray, mo: np.ndarray) -> np.ndarray:
    """lrs (jnum, nrs) row per vector; mo (n, norb) = Porb A2(norb,n) memory.
    Returns ChoT A3(norb, jnum, n) as a numpy array of shape (n, jnum, norb)."""
    m = np.ascontiguousarray(rs2lt, dtype=np.int64)
    L = _f64(lrs)
    C = _f64(mo)
    jnum, nrs = L.shape
    norb = C.shape[1]
    out = np.zeros((n, jnum, norb))
    lib().ora_cho_vtra2(n, nrs, _p(m, _i64p), jnum, _p(L), norb, _p(C), _p(out))
    return out


def cho_vtra0(n: int, rs2lt: np.ndarray, lrs: np.ndarray, mo: np.ndarray) -> np.ndarray:
    """As cho_vtra2 but A3(norb, n, jnum), returned with shape (jnum, n, norb)."""
    m = np.ascontiguousarray(rs2lt, dtype=np.int64)
    L = _f64(lrs)
    C = _f64(mo)
    jnum, nrs = L.shape
    norb = C.shape[1]
    out = np.zeros((jnum, n, norb))
    lib().ora_cho_vtra0(n, nrs, _p(m, _i64p), jnum, _p(L), norb, _p(C), _p(out))
    return out


def dgemm_tri_tn(A: np.ndarray, alpha: float = 1.0, beta: float = 0.0, c_lt=None) -> np.ndarray:
    """C_LT = beta C + alpha A^T A for A given as (n, k) memory = Fortran A(k,n)."""
    a = _f64(A)
    n, k = a.shape
    c = np.zeros(n * (n + 1) // 2) if c_lt is None else _f64(c_lt).copy()
    lib().ora_dgemm_tri_tn(n, k, alpha, _p(a), k, _p(a), k, beta, _p(c))
    return c


def cho_fscf(store: CholStore, flt: np.ndarray, nocc, porbs, dlt: np.ndarray, exfac: float,
             batch: int = 1 << 30) -> np.ndarray:
    """CHO_FSCF with nforb = 0, niorb = nocc.  flt (nden, npair) is updated in a
    copy which is returned; porbs[d] is (n, nocc[d]) (Porb A2(nocc,n) memory)."""
    nden = len(nocc)
    f = _f64(flt).reshape(nden, -1).copy()
    nf = np.zeros(nden, dtype=np.int32)
    ni = np.array(nocc, dtype=np.int32)
    cs = [_f64(p) for p in porbs]
    arr = (_dp * nden)(*[_p(c) for c in cs])
    d = _f64(dlt)
    rc = lib().ora_cho_fscf(ctypes.byref(store._c), nden, _p(f), _p(nf, _i32p), _p(ni, _i32p), arr,
                            _p(d), float(exfac), int(batch))
    if rc != 0:
        raise RuntimeError(f"ora_cho_fscf rc={rc}")
    return f


def cho_fock_dft_red(store: CholStore, dlt: np.ndarray, flt: np.ndarray, batch: int = 1 << 30):
    """CHO_FOCK_DFT_RED; returns the updated copy of flt (npair,)."""
    f = _f64(flt).reshape(-1).copy()
    rc = lib().ora_cho_fock_dft_red(ctypes.byref(store._c), _p(_f64(dlt)), _p(f), int(batch))
    if rc != 0:
        raise RuntimeError(f"ora_cho_fock_dft_red rc={rc}")
    return f


def cho_lk_scf(store: CholStore, flt: np.ndarray, nocc, porbs, plt: np.ndarray, factxi: float,
               nbassh=None, batch: int = 1 << 30):
    """CHO_LK_SCF without screening (nforb = 0, niorb = nocc).  porbs[d] is the
    (n, nocc[d]) AO x orbital matrix C[b, k].  Returns (flt, klt), both
    (nden, npair); klt in the shell-pair-blocked storage for shells nbassh
    (default: one shell)."""
    nden = len(nocc)
    n = store.n
    f = _f64(flt).reshape(nden, -1).copy()
    k = np.zeros_like(f)
    nf = np.zeros(nden, dtype=np.int32)
    ni = np.array(nocc, dtype=np.int32)
    cs = []
    for p in porbs:
        sq = np.zeros((n, n))
        sq[:, :p.shape[1]] = p
        cs.append(np.ascontiguousarray(sq.T))  # column-major A2(n,n): AO index fastest
    arr = (_dp * nden)(*[_p(c) for c in cs])
    sh = np.array([n] if nbassh is None else nbassh, dtype=np.int32)
    rc = lib().ora_cho_lk_scf(ctypes.byref(store._c), nden, _p(f), _p(k), _p(nf, _i32p), _p(ni, _i32p),
                              arr, _p(_f64(plt)), float(factxi), len(sh), _p(sh, _i32p), int(batch))
    if rc != 0:
        raise RuntimeError(f"ora_cho_lk_scf rc={rc}")
    return f, k


def cho_fmcscf(store: CholStore, flt: np.ndarray, nk: int, naorb: int, factxi: float,
               dlt: np.ndarray, do_active: bool, porb1, nchm: int, porb2, porb3, cmo,
               nfro: int, norb: int, batch: int = 1 << 30):
    """CHO_FMCSCF (nforb = 0, niorb = nk).  Returns (flt (2,npair), w_pwxy)."""
    f = _f64(flt).reshape(2, -1).copy()
    d = _f64(dlt).reshape(2, -1)
    n = store.n
    p1 = _f64(porb1) if nk else np.zeros(1)
    p2 = _f64(porb2) if nchm else np.zeros(1)
    p3 = _f64(porb3) if naorb else np.zeros(1)
    cm = _f64(np.asarray(cmo).T) if cmo is not None else np.zeros(1)  # column-major CMO(n,n)
    nxy = naorb * (naorb + 1) // 2
    w = np.zeros(max(norb * naorb * nxy, 1))
    rc = lib().ora_cho_fmcscf(ctypes.byref(store._c), _p(f), 0, nk, naorb, float(factxi), _p(d),
                              int(do_active), _p(p1), nchm, _p(p2), _p(p3), _p(w), _p(cm), nfro,
                              norb, 1.0, int(batch))
    if rc != 0:
        raise RuntimeError(f"ora_cho_fmcscf rc={rc}")
    return f, w[:norb * naorb * nxy]


def cho_lk_casscf(store: CholStore, dlt: np.ndarray, flt: np.ndarray, msq, factxi: float,
                  nfiorb: int, naorb: int, nchm: int, ash1, ash2, do_active: bool, cmo,
                  nfro: int, norb: int, exfac: float, batch: int = 1 << 30):
    """CHO_LK_CASSCF without screening.  msq (n, >= nfiorb) and ash2 (n, >= nchm)
    are AO x vector matrices, ash1 (n, naorb) the active MOs.  Returns (flt, w)."""
    n = store.n

    def square(a, k):
        sq = np.zeros((n, n))
        if a is not None and k > 0:
            sq[:, :k] = np.asarray(a)[:, :k]
        return np.ascontiguousarray(sq.T)  # column-major A2(n,n)

    f = _f64(flt).reshape(2, -1).copy()
    d = _f64(dlt).reshape(2, -1)
    a1 = _f64(ash1) if naorb else np.zeros(1)
    cm = _f64(np.asarray(cmo).T) if cmo is not None else np.zeros(1)
    nxy = naorb * (naorb + 1) // 2
    w = np.zeros(max(norb * naorb * nxy, 1))
    rc = lib().ora_cho_lk_casscf(ctypes.byref(store._c), _p(d), _p(f), _p(square(msq, nfiorb)), _p(w),
                                 float(factxi), nfiorb, naorb, nchm, _p(a1), _p(square(ash2, nchm)),
                                 int(do_active), _p(cm), nfro, norb, float(exfac), int(batch))
    if rc != 0:
        raise RuntimeError(f"ora_cho_lk_casscf rc={rc}")
    return f, w[:norb * naorb * nxy]


def lk_thresholds(thrcom: float, dfmat: float, dmpk: float, nocc, numvt: int, estimate: bool = False,
                  maxred: int = 1):
    """tau and thrv per density (cho_lk_scf.F90:143-171)."""
    nd = len(nocc)
    no = np.array(nocc, dtype=np.int32)
    tau, thrv = np.zeros(nd), np.zeros(nd)
    lib().ora_lk_thresholds(float(thrcom), float(dfmat), float(dmpk), nd, _p(no, _i32p), int(numvt),
                            int(bool(estimate)), int(maxred), _p(tau), _p(thrv))
    return tau, thrv


def _lk_struct(store, nbassh, diag, thrcom, maxred, estimate, update, keep):
    sh = np.ascontiguousarray(nbassh, dtype=np.int32)
    dg = _f64(diag).reshape(-1)
    keep += [sh, dg]
    return _LK(len(sh), _p(sh, _i32p), _p(dg), float(thrcom),
               int(store._c.nred if maxred is None else maxred), int(bool(estimate)), int(bool(update)), 0)


def cho_lk_scf_screened(store: CholStore, flt: np.ndarray, nocc, porbs, plt: np.ndarray, factxi: float,
                        nbassh, diag, thrcom: float, nscreen: int = 10, dmpk: float = 0.1,
                        dfmat: float = 1.0, maxred=None, estimate: bool = False, update: bool = True,
                        batch: int = 25):
    """CHO_LK_SCF with the reference's LK screening (nforb = 0, niorb = nocc).
    porbs[d]: (n, nocc[d]) AO x orbital matrix; diag: the integral diagonal
    on the LT (CHO_IODIAG); batch: Cho_LK_MaxVecPerBatch.  Returns (flt, klt,
    stats) with stats = [listed (batch, orbital, shell), half-transformed
    shell pairs, K blocks]."""
    nden = len(nocc)
    n = store.n
    f = _f64(flt).reshape(nden, -1).copy()
    k = np.zeros_like(f)
    nf = np.zeros(nden, dtype=np.int32)
    ni = np.array(nocc, dtype=np.int32)
    cs = []
    for p in porbs:
        sq = np.zeros((n, n))
        sq[:, :p.shape[1]] = p
        cs.append(np.ascontiguousarray(sq.T))
    arr = (_dp * nden)(*[_p(c) for c in cs])
    keep = []
    lk = _lk_struct(store, nbassh, diag, thrcom, maxred, estimate, update, keep)
    stats = np.zeros(3, dtype=np.int64)
    rc = lib().ora_cho_lk_scf_scr(ctypes.byref(store._c), nden, _p(f), _p(k), _p(nf, _i32p), _p(ni, _i32p),
                                  arr, _p(_f64(plt)), float(factxi), int(nscreen), float(dmpk), float(dfmat),
                                  ctypes.byref(lk), int(batch), _p(stats, _i64p))
    if rc != 0:
        raise RuntimeError(f"ora_cho_lk_scf_scr rc={rc}")
    return f, k, stats


def cho_lk_casscf_screened(store: CholStore, dlt: np.ndarray, flt: np.ndarray, msq, factxi: float,
                           nfiorb: int, naorb: int, nchm: int, ash1, ash2, do_active: bool, cmo,
                           nfro: int, norb: int, exfac: float, nbassh, diag, thrcom: float,
                           nscreen: int = 10, dmpk: float = 0.1, dfmat: float = 1.0, maxred=None,
                           estimate: bool = False, update: bool = True, batch: int = 25):
    """CHO_LK_CASSCF with the reference's LK screening.  Arguments as
    cho_lk_casscf plus the LK state.  Returns (flt, w, stats)."""
    n = store.n

    def square(a, k):
        sq = np.zeros((n, n))
        if a is not None and k > 0:
            sq[:, :k] = np.asarray(a)[:, :k]
        return np.ascontiguousarray(sq.T)

    f = _f64(flt).reshape(2, -1).copy()
    d = _f64(dlt).reshape(2, -1)
    a1 = _f64(ash1) if naorb else np.zeros(1)
    cm = _f64(np.asarray(cmo).T) if cmo is not None else np.zeros(1)
    nxy = naorb * (naorb + 1) // 2
    w = np.zeros(max(norb * naorb * nxy, 1))
    keep = []
    lk = _lk_struct(store, nbassh, diag, thrcom, maxred, estimate, update, keep)
    stats = np.zeros(3, dtype=np.int64)
    rc = lib().ora_cho_lk_casscf_scr(ctypes.byref(store._c), _p(d), _p(f), _p(square(msq, nfiorb)), _p(w),
                                     float(factxi), nfiorb, naorb, nchm, _p(a1), _p(square(ash2, nchm)),
                                     int(do_active), int(nscreen), float(dmpk), float(dfmat), _p(cm), nfro,
                                     norb, float(exfac), ctypes.byref(lk), int(batch), _p(stats, _i64p))
    if rc != 0:
        raise RuntimeError(f"ora_cho_lk_casscf_scr rc={rc}")
    return f, w[:norb * naorb * nxy], stats


def cd_incore(x: np.ndarray, thr: float, mxvec=None):
    """Returns (rc, V (n, numcho)); x is a symmetric square matrix."""
    a = _f64(x).copy()
    n = a.shape[0]
    mx = n if mxvec is None else int(mxvec)
    vec = np.zeros((max(mx, 1), n))
    num = np.zeros(1, dtype=np.int32)
    rc = lib().ora_cd_incore(_p(a), n, _p(vec), mx, _p(num, _i32p), float(thr))
    return int(rc), vec[:num[0]].T.copy()


def energy(h: np.ndarray, dlt: np.ndarray, g: np.ndarray, enuc: float = 0.0) -> float:
    h, d, g = _f64(h), _f64(dlt), _f64(g)
    return float(lib().ora_energy(len(d), _p(h), _p(d), _p(g), float(enuc)))
