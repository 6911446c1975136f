"""Precision ladder for the exchange build.

The exchange energy is a sum over Cholesky vectors, E_K = -f sum_J
||C^T L^J C||_F^2, and the contribution of vector J is bounded by its weight
w_J = ||L^J||_F^2 (for orthonormal C).  Cholesky vectors come in pivot order,
so the weights decay and a low precision applied to the tail [j, nchol) costs
an energy error that scales with the tail sum T(j) = sum_{J >= j} w_J:

    |dE_mode(j)| ~= kappa_mode * T(j)

kappa_mode is measured once per system (``calibrate``: one exchange build in
each precision against FP64); ``plan`` then places the split points so each
reduced-precision segment stays below a share of the energy tolerance, and
``mjk_exchange_mixed`` runs [0, j_fp32) in FP64, [j_fp32, j_bf16) in FP32 and
[j_bf16, nchol) in split-BF16.  Rounding errors of a full-precision-mode build
partly cancel, so the linear model can under-predict a tail's error;
``tune`` therefore builds the planned ladder, measures its energy error
against FP64 and tightens the budget until the measured error is within the
acceptance share of the tolerance.  ``tune`` also times one build in each
reduced mode and leaves split-BF16 out when it is not faster than FP32 on
this system (FP32 is the more accurate of the two, so it then takes the whole
reduced tail).

The reference has no mixed-precision exchange: src/fock_util/cho_fscf.F90:
250,276 runs every vector in FP64.
"""
from __future__ import annotations

import time
from dataclasses import dataclass

import numpy as np
import torch


@dataclass
class Ladder:
    j_fp32: int
    j_bf16: int
    predicted: float  # predicted |dE| of the two reduced-precision segments
    measured: float = float("nan")  # |dE| of this ladder against FP64, when verified
    nchol: int = -1  # number of vectors the split points refer to


def tail_sums(w) -> np.ndarray:
    """T[j] = sum_{J >= j} w_J, length nchol + 1 (T[nchol] = 0)."""
    w = np.asarray(w.detach().cpu() if isinstance(w, torch.Tensor) else w, dtype=np.float64)
    t = np.zeros(w.size + 1)
    t[:-1] = np.cumsum(w[::-1])[::-1]
    return t


def plan(w, kappa_fp32: float, kappa_bf16: float, tol: float = 1e-6,
         share: float = 0.25) -> Ladder:
    """Smallest split points whose predicted energy error per reduced segment
    is at most share * tol (FP32 segment [j_fp32, j_bf16), split-BF16 segment
    [j_bf16, nchol))."""
    t = tail_sums(w)
    m = t.size - 1
    budget = share * tol
    jb = m
    while jb > 0 and kappa_bf16 * t[jb - 1] <= budget:
        jb -= 1
    j3 = jb
    while j3 > 0 and kappa_fp32 * (t[j3 - 1] - t[jb]) <= budget:
        j3 -= 1
    pred = kappa_fp32 * (t[j3] - t[jb]) + (kappa_bf16 * t[jb] if jb < m else 0.0)
    return Ladder(int(j3), int(jb), float(pred), nchol=int(m))


def speedup_bound(plans, work=None, rate_fp32: float = float("inf"),
                  rate_bf16: float = float("inf")) -> float:
    """Exchange speedup over FP64 of ladders whose segments run at the given
    per-vector rates relative to FP64: 1 / (f64 + f32 / rate_fp32 + fb /
    rate_bf16), with f the shares of the exchange work in each mode.  The
    work of one vector of a system is the same for every vector (2 n^2 o +
    n (n + 1) o flop), so a system's shares are its vector counts; ``work``
    weights the systems of a batch (per-vector work of each, default 1).
    With the default infinite rates this is the cap the FP64 head sets:
    even free reduced-precision segments give no more than 1 / f64."""
    plans = [plans] if isinstance(plans, Ladder) else list(plans)
    work = [1.0] * len(plans) if work is None else [float(x) for x in work]
    nchol = [q.nchol for q in plans]
    if any(m < 0 for m in nchol):
        raise ValueError("speedup_bound needs ladders that carry nchol (plan/tune set it)")
    tot = sum(w * m for w, m in zip(work, nchol))
    f64 = sum(w * q.j_fp32 for w, q in zip(work, plans)) / tot
    f32 = sum(w * (q.j_bf16 - q.j_fp32) for w, q in zip(work, plans)) / tot
    fb = sum(w * (m - q.j_bf16) for w, q, m in zip(work, plans, nchol)) / tot
    t = f64 + (f32 / rate_fp32 if f32 > 0 else 0.0) + (fb / rate_bf16 if fb > 0 else 0.0)
    return float("inf") if t == 0.0 else 1.0 / t


def exchange_energy_error(k_test: torch.Tensor, k_ref: torch.Tensor, dlt: torch.Tensor,
                          factor: float = 0.5, per_density: bool = False) -> float:
    """|factor * sum_d sum_LT (K_test - K_ref)_d * dlt_d| for folded densities;
    per_density: the largest of the densities' own |factor * sum_LT (...)_d|
    (the bound every state of a state-averaged batch must meet)."""
    prod = (k_test - k_ref) * dlt
    if per_density:
        rows = prod.reshape(-1, prod.shape[-1])
        return float(torch.max(torch.abs(factor * torch.sum(rows, dim=1))).item())
    return abs(factor * torch.sum(prod).item())


def calibrate(jk, corbs, dlt: torch.Tensor, factor: float = 0.5, w=None, per_density: bool = False):
    """Measure kappa_fp32 and kappa_bf16 for this system: one exchange build
    per precision against FP64.  dlt holds the folded densities that weight
    K in the energy (one row per orbital set) and factor the exchange energy
    factor (1/2 for the closed-shell E = h.D + 1/2 (J - K).D with D = 2 C C^T).
    Returns (kappa_fp32, kappa_bf16, w, k_fp64)."""
    if w is None:
        w = jk.vector_norms()
    total = float(torch.sum(w).item())
    k64 = jk.exchange(corbs, precision="fp64")
    kappas = []
    for prec in ("fp32", "bf16x3"):
        k = jk.exchange(corbs, precision=prec)
        de = exchange_energy_error(k, k64, dlt, factor, per_density)
        kappas.append(max(de, 1e-300) / max(total, 1e-300))
    return kappas[0], kappas[1], w, k64


def mode_times(jk, corbs, repeats: int = 2):
    """Seconds per exchange build in FP32 and in split-BF16 on this system."""
    out = []
    for prec in ("fp32", "bf16x3"):
        jk.exchange(corbs, precision=prec)
        jk.sync()
        t0 = time.perf_counter()
        for _ in range(repeats):
            jk.exchange(corbs, precision=prec)
        jk.sync()
        out.append((time.perf_counter() - t0) / repeats)
    return out[0], out[1]


def tune(jk, corbs, dlt: torch.Tensor, factor: float = 0.5, tol: float = 1e-6,
         share: float = 0.25, accept: float = 0.5, rounds: int = 6,
         per_density: bool = False) -> Ladder:
    """Calibrate, plan and verify: returns a ladder whose measured |dE|
    against the FP64 build is at most accept * tol (FP64 throughout if no
    ladder within ``rounds`` tightening steps qualifies).  per_density: |dE|
    is the largest error of any one density (orbital set) instead of the
    error of their sum."""
    k32, kb3, w, k64 = calibrate(jk, corbs, dlt, factor, per_density=per_density)
    t32, tb3 = mode_times(jk, corbs)
    if tb3 >= t32:
        kb3 = float("inf")  # no split-BF16 segment: FP32 is as fast and more accurate
    m = int(w.numel())
    s = share
    for _ in range(rounds):
        p = plan(w, k32, kb3, tol, s)
        if p.j_fp32 >= m:
            break
        k = jk.exchange_mixed(corbs, p.j_fp32, p.j_bf16)
        de = exchange_energy_error(k, k64, dlt, factor, per_density)
        p.measured = de
        if de <= accept * tol:
            return p
        s *= 0.5 * accept * tol / max(de, 1e-300)
    return Ladder(m, m, 0.0, 0.0, nchol=m)
