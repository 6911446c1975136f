#!/usr/bin/env python3
"""Benchmark of the Cholesky J/K Fock build on one MI355X.

Default workload (the largest single-GPU configuration): 64 independent
systems, nbas ~ U[100,600], nchol = round(k nbas) with k ~ U[4,6], nocc =
10-20% of nbas (seeded), Cholesky vectors resident in HBM.  One "step" is one
replay of the captured HIP graph that runs, for every system, the closed-shell
energy build (density, J, K, Fock, E) and the hand-written backward (dE/dC,
dE/dL).  value = systems built (forward + backward) per second in FP64.  The
same line reports the forward-only replay, the precision ladder (FP64 / FP32 /
split-BF16 split points tuned per system for |dE| <= 1e-6 Eh) with its rate
and errors, the roofline of the dominant kernel from profiled eager steps of
the same batch, the CPU oracle on a sample of the same workload, and, as
secondary objects, the single n=512 build and the n=1200 build.
Prints ONE JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3, "bf16x3": 2500.0 / 3.0}


def progress(msg):
    """One progress line on stderr per finished phase (the JSON line stays alone on stdout)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--workload", default="batch64", choices=["batch64", "single512"],
                   help="headline workload: the 64-system batch (default) or the single n=512 build")
    p.add_argument("--no-secondary", action="store_true",
                   help="skip the secondary n=512 and n=1200 objects")
    p.add_argument("--nbas", type=int, default=512)
    p.add_argument("--nchol", type=int, default=2048)
    p.add_argument("--nocc", type=int, default=64)
    p.add_argument("--precision", default="fp64", choices=["fp64", "fp32", "bf16x3"])
    p.add_argument("--seed", type=int, default=2024)
    p.add_argument("--two-pass", action="store_true",
                   help="FP64: run the Coulomb V pass on its own instead of inside the half-transform")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--no-mixed", action="store_true")
    p.add_argument("--extra", action="store_true",
                   help="also time the 9-density state-averaged batch; adds an 'extra' object")
    p.add_argument("--copy-probe", action="store_true",
                   help="after the timed region, copy 1 GiB device to device (calibrates HBM counters)")
    p.add_argument("--cpu-sample-vectors", type=int, default=1 << 30)
    p.add_argument("--kernel-profile", action="store_true",
                   help="64-system batch only: set up, autotune, one graph replay, then (after a 0.5 s "
                        "pause that marks it in a kernel trace) one FP64 forward + backward step with the "
                        "systems one after another; prints that step's per-phase kernel times")
    p.add_argument("--profile-precision", choices=["fp64", "ladder"], default="fp64",
                   help="with --kernel-profile: the precision of the profiled step (ladder: split points "
                        "tuned per system for |dE| <= 1e-6 Eh first)")
    return p.parse_args()


class _Route:
    """One way of running the oracle's DGEMV calls for timing: its plain
    OpenMP loops, or the host BLAS (as the reference links one) with the BLAS
    threads capped at `threads`.  Restores the plain loops on exit."""

    def __init__(self, threads, blas):
        self.threads, self.use = int(threads), bool(blas)

    def __enter__(self):
        import oracle
        from threadpoolctl import threadpool_limits

        oracle.set_threads(self.threads)
        self.limits = threadpool_limits(self.threads, "blas")
        self.name = oracle.use_blas(True) if self.use else "plain OpenMP loops"
        return self

    def __exit__(self, *exc):
        import oracle

        oracle.use_blas(False)
        self.limits.restore_original_limits()
        return False


def fastest_route(threads, run):
    """Times run() with the plain loops and with the host BLAS on `threads`
    threads; returns (seconds, route name) of the faster."""
    best = None
    for blas in (False, True):
        with _Route(threads, blas) as r:
            t = run()
        if best is None or t < best[0]:
            best = (t, r.name)
    return best


def host_threads():
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


def cpu_baseline(n, M, o, seed, sample, repeats=5):
    """The CPU oracle (C restatement of CHO_FSCF, OpenMP over the host cores,
    its DGEMVs as plain loops or through the host BLAS, whichever is faster)
    on the same workload; `sample` vectors (all by default), median of
    `repeats` builds, scaled to the full nchol when sampled.  The parity
    reference is the plain-loop oracle."""
    import numpy as np

    import oracle
    from openmolcas_amd.synthetic import chol_vectors_np, fold, npair, orbitals_np

    ms = min(sample, M)
    L = chol_vectors_np(n, ms, seed)
    C = orbitals_np(n, o, seed + 1)
    dlt = fold(C @ C.T)
    st = oracle.CholStore.full(L)
    flt = np.zeros((1, npair(n)))
    want = oracle.cho_fscf(st, flt, [o], [C], dlt, 0.5, batch=256)  # the parity reference
    ref = (L, C, dlt, want)
    threads = host_threads()

    def run():
        oracle.cho_fscf(st, flt, [o], [C], dlt, 0.5, batch=256)  # warm
        times = []
        for _ in range(repeats):
            t0 = time.perf_counter()
            oracle.cho_fscf(st, flt, [o], [C], dlt, 0.5, batch=256)
            times.append(time.perf_counter() - t0)
        return float(np.median(times))

    dt, route = fastest_route(threads, run)
    return {
        "value": round(1.0 / (dt * M / ms), 6),
        "unit": "builds/s",
        "cores": threads,
        "kind": "port",
        "dgemv": route,
        "sample": f"oracle CHO_FSCF restatement (C, OpenMP, {threads} threads, DGEMVs: {route}, the faster of "
                  f"the plain loops and the host BLAS) on {ms} of {M} vectors, nbas={n} nocc={o}, batch 256 vectors, "
                  f"median of {repeats}: {dt:.2f} s per build" + (f", scaled x{M / ms:.1f}" if ms < M else ""),
    }, ref


def gpu_parity(ref, dev):
    """The FP64 drop-in (mjk_cho_fscf) on the CPU baseline's sample against the
    oracle's FLT of the same inputs: max-abs error of F = J[D] - 1/2 K[D]."""
    import numpy as np

    from openmolcas_amd.fock import JKContext

    L, C, dlt, want = ref
    jk = JKContext(C.shape[0], L.shape[0], device=dev)
    jk.upload_packed(L)
    got = np.zeros_like(want)
    jk.cho_fscf(1, got, [0], [C.shape[1]], [C], dlt, 0.5)
    out = {"against": "oracle CHO_FSCF FLT on the cpu_baseline sample (FP64 drop-in)",
           "flt_maxabs_err": float(np.max(np.abs(got - want))),
           "flt_maxabs": float(np.max(np.abs(want)))}
    # the drop-in as a Fortran caller sees it: host DLT, Porb and FLT copied in,
    # FLT copied back, synchronised (PCIe-inclusive; the vectors stay in HBM)
    reps = 10
    flt = np.zeros_like(want)
    t0 = time.perf_counter()
    for _ in range(reps):
        jk.cho_fscf(1, flt, [0], [C.shape[1]], [C], dlt, 0.5)
    dt = (time.perf_counter() - t0) / reps
    jk.close()
    out["dropin_host_buffers_ms"] = round(1e3 * dt, 4)
    out["dropin_host_buffers_builds_per_s"] = round(1.0 / dt, 2)
    return out


def energy_of(ctx, dlt, h, J, K):
    """E = tr(D h) + 1/2 tr(D J) - 1/4 tr(D K) through the FP64 energy step:
    G = J - 1/2 K (LT), E = sum h.D + 1/2 sum G.D."""
    import torch

    G = torch.zeros_like(J)
    ctx.fock_update(G, J, [0], [1.0], K, [-0.5])
    return ctx.energy(h, dlt, G).item()


def mixed_vs_fp64(ctx, dlt, h, C, n, M, o, ms_fp64, args):
    """Times the exchange alone in FP64 and in each mixed mode (same inputs) and
    reports speedups and |dE| against the FP64 build."""
    import torch

    J = ctx.coulomb(dlt)
    res = {}
    K64 = ctx.exchange([C], precision="fp64")
    e64 = energy_of(ctx, dlt, h, J, K64)
    tk = {}
    for prec in ("fp64", "fp32", "bf16x3"):
        try:
            ctx.autotune([C], prec)  # each mode timed with its own tiles
            ctx.exchange([C], precision=prec)
        except Exception as exc:  # mode not built
            res[prec] = {"error": str(exc)}
            continue
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            K = ctx.exchange([C], precision=prec)
        torch.cuda.synchronize()
        tk[prec] = (time.perf_counter() - t0) / args.steps
        e = energy_of(ctx, dlt, h, J, K)
        res[prec] = {"exchange_ms": round(1e3 * tk[prec], 4), "E": e, "dE": abs(e - e64),
                     "K_maxabs_err": float(torch.max(torch.abs(K - K64)).item())}
    fk = M * o * (2.0 * n * n + n * (n + 1))
    for prec in tk:
        res[prec]["exchange_TFs"] = round(fk / tk[prec] / 1e12, 2)
        if prec != "fp64":
            res[prec]["exchange_speedup_vs_fp64"] = round(tk["fp64"] / tk[prec], 3)
    if "fp64" in tk:  # the plain FP64 exchange (mjk_exchange, no V work) against the MFMA peak
        res["fp64"]["exchange_frac_peak"] = round(fk / tk["fp64"] / 1e12 / PEAK_TFLOPS["fp64"], 4)
    if "fp32" in tk and "bf16x3" in tk:
        ctx.autotune([C], "fp64")  # the ladder's head is FP64
        res["ladder"] = ladder_vs_fp64(ctx, dlt, h, C, J, K64, e64, tk["fp64"], res, args.steps)
    return res


def ladder_vs_fp64(ctx, dlt, h, C, J, K64, e64, t64, res, steps):
    """Precision ladder: split points calibrated, planned and verified for
    |dE| <= 1e-6 Eh (openmolcas_amd.ladder.tune), then the ladder build timed
    and its energy checked against FP64 with the same energy expression."""
    import torch

    from openmolcas_amd import ladder

    # energy_of uses G = J - 1/2 K and E = h.D + 1/2 G.D: dE = 1/4 dK.D
    plan = ladder.tune(ctx, [C], dlt, factor=0.25, tol=1e-6)
    K = ctx.exchange_mixed([C], plan.j_fp32, plan.j_bf16)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        K = ctx.exchange_mixed([C], plan.j_fp32, plan.j_bf16)
    torch.cuda.synchronize()
    t = (time.perf_counter() - t0) / steps
    e = energy_of(ctx, dlt, h, J, K)
    r32 = t64 / (res["fp32"]["exchange_ms"] * 1e-3)
    rb3 = t64 / (res["bf16x3"]["exchange_ms"] * 1e-3)
    return {"j_fp32": plan.j_fp32, "j_bf16": plan.j_bf16, "nchol": int(ctx.nchol),
            "dE": abs(e - e64), "exchange_ms": round(1e3 * t, 4),
            "exchange_speedup_vs_fp64": round(t64 / t, 3),
            # the FP64 head's cap (free reduced segments) and the Amdahl bound at
            # the measured whole-build FP32 / split-BF16 rates
            "speedup_cap_fp64_head": round(ladder.speedup_bound(plan), 3),
            "speedup_bound_at_mode_rates": round(ladder.speedup_bound(plan, None, r32, rb3), 3)}


def batch_traffic(family, launches):
    """HBM bytes per launch of a kernel family in the 64-system step, from the
    committed rocprofv3 summary of that step (FETCH_SIZE + WRITE_SIZE summed
    over the step, calibrated), or None."""
    path = os.path.join(ROOT, "profiles", "batch64_fp64_fwd_bwd.md")
    if not os.path.exists(path) or not launches:
        return None
    try:
        data = json.loads(open(path).read().split("```json", 1)[1].split("```", 1)[0])
    except (IndexError, ValueError):
        return None
    v = data.get(family)
    return None if v is None else (v["fetch_bytes"] + v["write_bytes"]) / launches


def profiled_traffic(kernel_prefix, n, M, o, precision):
    """HBM bytes per dispatch of `kernel_prefix` from the committed rocprofv3
    summary of this workload (FETCH_SIZE + WRITE_SIZE, calibrated), or None."""
    path = os.path.join(ROOT, "profiles", f"jk_build_n{n}_m{M}_o{o}_{precision}.md")
    if not os.path.exists(path):
        return None
    text = open(path).read()
    try:
        blob = text.split("```json", 1)[1].split("```", 1)[0]
        data = json.loads(blob)
    except (IndexError, ValueError):
        return None
    for k, v in data.items():
        if k.startswith(kernel_prefix):
            return v["fetch_bytes"] + v["write_bytes"]
    return None


def _timeit(fn, reps):
    import torch

    fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / reps


def build1200(dev):
    """Secondary: the n=1200/nchol=6000/nocc=150 build (P = 2CC^T) with its
    backward, FP64, FP32, split-BF16 and the tuned ladder (|dE| <= 1e-6 Eh)."""
    import torch

    from openmolcas_amd import ladder
    from openmolcas_amd.fock import JKContext
    from openmolcas_amd.synthetic import chol_vectors_torch, npair, orbitals_torch

    n, M, o = 1200, 6000, 150
    jk = JKContext(n, M, device=dev)
    L = jk.device_vectors()
    chol_vectors_torch(n, M, 3, device=dev, out=L)
    C = orbitals_torch(n, o, 4, device=dev)
    dlt = jk.density(C, 2.0)
    res = {"workload": "nbas=1200 nchol=6000 nocc=150, P = 2CC^T (secondary)"}
    J, V = jk.coulomb(dlt, want_v=True)
    K64 = jk.exchange([C])
    G = torch.zeros_like(J)
    jk.fock_update(G, J, [0], [1.0], K64, [-1.0])
    h = torch.zeros_like(J)
    e64 = jk.energy(h, dlt, G).item()
    res["coulomb_ms"] = round(1e3 * _timeit(lambda: jk.coulomb(dlt), 3), 3)
    res["coulomb_GBs"] = round(2.0 * M * npair(n) * 8 / (res["coulomb_ms"] * 1e-3) / 1e9, 1)
    jk.autotune([C], "fp64")
    res["builds_per_s_fp64"] = round(1.0 / _timeit(lambda: jk.coulomb_exchange(dlt, [C]), 3), 3)
    fk = M * o * (2.0 * n * n + n * (n + 1))
    dC64 = jk.grad_orbitals(C, G, 4.0)
    for prec in ("fp64", "fp32", "bf16x3"):
        jk.autotune([C], prec)  # each mode timed with its own tiles
        t = _timeit(lambda: jk.exchange([C], precision=prec), 2)
        K = jk.exchange([C], precision=prec)
        Gp = torch.zeros_like(J)
        jk.fock_update(Gp, J, [0], [1.0], K, [-1.0])
        res[f"exchange_{prec}_ms"] = round(1e3 * t, 3)
        res[f"exchange_{prec}_TFs"] = round(fk / t / 1e12, 2)
        res[f"dE_{prec}"] = abs(jk.energy(h, dlt, Gp).item() - e64)
        if prec != "fp64":
            res[f"dEdC_maxabs_err_{prec}"] = float(torch.max(torch.abs(jk.grad_orbitals(C, Gp, 4.0) - dC64)).item())
    jk.autotune([C], "fp64")  # the ladder's head is FP64
    plan = ladder.tune(jk, [C], dlt, factor=0.5, tol=1e-6)  # G = J - K[C]: dE = 1/2 dK.D
    t = _timeit(lambda: jk.exchange_mixed([C], plan.j_fp32, plan.j_bf16), 2)
    K = jk.exchange_mixed([C], plan.j_fp32, plan.j_bf16)
    Gp = torch.zeros_like(J)
    jk.fock_update(Gp, J, [0], [1.0], K, [-1.0])
    res["ladder"] = {"j_fp32": plan.j_fp32, "j_bf16": plan.j_bf16, "exchange_ms": round(1e3 * t, 3),
                     "dE": abs(jk.energy(h, dlt, Gp).item() - e64),
                     "dEdC_maxabs_err": float(torch.max(torch.abs(jk.grad_orbitals(C, Gp, 4.0) - dC64)).item()),
                     "speedup_vs_fp64": round(res["exchange_fp64_ms"] / (1e3 * t), 3)}
    dl = torch.empty((M, npair(n)), dtype=torch.float64, device=dev)
    res["backward_dC_ms"] = round(1e3 * _timeit(lambda: jk.grad_orbitals(C, G, 4.0), 3), 3)
    res["backward_dL_ms"] = round(1e3 * _timeit(
        lambda: jk.grad_vectors(C, V[:, 0].contiguous(), 1.0, -2.0, out=dl), 2), 3)
    del dl, L
    jk.close()
    torch.cuda.empty_cache()
    return res


def state_averaged(dev):
    """Extra: the 9-density state-averaged batch (n=800, nchol=3600)."""
    import numpy as np
    import torch

    from openmolcas_amd.fock import JKContext
    from openmolcas_amd.synthetic import chol_vectors_torch, npair, orbitals_torch

    n, M = 800, 3600
    rng = np.random.default_rng(11)
    ranks = [120] + [int(r) for r in rng.integers(4, 17, size=8)]
    jk = JKContext(n, M, device=dev)
    L = jk.device_vectors()
    chol_vectors_torch(n, M, 12, device=dev, out=L)
    Q = orbitals_torch(n, sum(ranks), 13, device=dev)
    Cs, dl, o0 = [], [], 0
    ii, jj = torch.tril_indices(n, n, device=dev)
    w = torch.where(ii == jj, 1.0, 2.0).to(torch.float64)
    for d, r in enumerate(ranks):
        occ = torch.full((r,), 2.0, dtype=torch.float64, device=dev) if d == 0 else \
            torch.tensor(rng.uniform(0.05, 1.95, r), dtype=torch.float64, device=dev)
        Cd = (Q[:, o0:o0 + r] * torch.sqrt(occ)).contiguous()
        o0 += r
        Cs.append(Cd)
        dl.append(((Cd @ Cd.T)[ii, jj] * w).contiguous())
    D = torch.stack(dl)
    res = {"workload": f"nbas=800 nchol=3600, 9 densities ranks {ranks}"}
    tj = _timeit(lambda: jk.coulomb(D), 3)
    res["coulomb_9dens_ms"] = round(1e3 * tj, 3)
    res["coulomb_GBs"] = round(2.0 * M * npair(n) * 8 / tj / 1e9, 1)
    fk = sum(M * r * (2.0 * n * n + n * (n + 1)) for r in ranks)
    for prec in ("fp64", "fp32", "bf16x3"):
        t = _timeit(lambda: jk.exchange(Cs, precision=prec), 3)
        res[f"exchange_{prec}_ms"] = round(1e3 * t, 3)
        res[f"exchange_{prec}_TFs"] = round(fk / t / 1e12, 2)
    res["builds_per_s_fp64"] = round(1.0 / (tj + res["exchange_fp64_ms"] * 1e-3), 3)
    del L
    jk.close()
    torch.cuda.empty_cache()
    return res


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def batch_cpu_baseline(shapes, seed, frac=32):
    """The CPU oracle (C restatement of CHO_FSCF, its DGEMVs as plain loops or
    through the host BLAS, whichever is faster) on the batch workload: every
    system's forward J/K build with 1/frac of its vectors (at least 8), on
    all host cores, then the first 16 systems on one core; both scaled to the
    full vector counts (the build is linear in the vector count).  The
    reference has no backward, so the CPU rate counts forward builds."""
    import numpy as np

    import oracle
    from openmolcas_amd.synthetic import chol_vectors_np, fold, npair, orbitals_np

    def run(systems):
        tot = 0.0
        for i, (n, M, o) in systems:
            ms = max(8, M // frac)
            L = chol_vectors_np(n, ms, seed + 7 * i)
            C = orbitals_np(n, o, seed + 7 * i + 1)
            dlt = fold(2.0 * C @ C.T)
            st = oracle.CholStore.full(L)
            flt = np.zeros((1, npair(n)))
            t0 = time.perf_counter()
            oracle.cho_fscf(st, flt, [o], [np.sqrt(2.0) * C], dlt, 0.5, batch=256)
            tot += (time.perf_counter() - t0) * M / ms
        return tot

    threads = host_threads()
    t_all, route = fastest_route(threads, lambda: run(list(enumerate(shapes))))
    sub = list(enumerate(shapes))[:16]
    t_sub, route1 = fastest_route(1, lambda: run(sub))
    t_one = t_sub * sum(M * o * n * n for n, M, o in shapes) / sum(M * o * n * n for _, (n, M, o) in sub)
    oracle.set_threads(threads)
    return {
        "value": round(len(shapes) / t_all, 4),
        "unit": "builds/s",
        "cores": threads,
        "kind": "port",
        "dgemv": route,
        "sample": f"oracle CHO_FSCF restatement (C, OpenMP, DGEMVs: {route}, the faster of the plain loops "
                  f"and the host BLAS) on every one of the {len(shapes)} systems with 1/{frac} of its vectors (>= 8), forward "
                  f"J/K only (the reference has no backward), scaled to the full vector counts: {t_all:.1f} s "
                  f"per batch on {threads} threads",
        "single_core": {"value": round(len(shapes) / t_one, 5), "unit": "builds/s", "cores": 1,
                        "dgemv": route1,
                        "sample": f"the first 16 systems at 1/{frac} of their vectors, scaled by the work "
                                  f"M o n^2 of all {len(shapes)}"},
        "cpu_model": cpu_model(),
    }


def batch_workload(args, dev):
    """The 64-system uneven batch: FP64 forward + backward graph replays (the
    timed region), the forward-only replay, the tuned ladder, and a profiled
    eager step for the per-kernel roofline."""
    import torch

    from openmolcas_amd import ladder
    from openmolcas_amd.batch import MoleculeBatch, uneven_shapes

    shapes = uneven_shapes(64, seed=5)
    batch = MoleculeBatch(shapes, device=dev, seed=3)
    progress(f"batch of {len(shapes)} systems set up")
    batch.autotune("fp64")
    progress("autotune done")
    nsys = len(shapes)

    def timed(precision, backward, steps, warmup):
        batch.capture(precision, backward)
        for _ in range(warmup):
            batch.replay()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            batch.replay()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps

    t64 = timed("fp64", True, args.steps, args.warmup)
    progress(f"fp64 forward+backward replays: {1e3 * t64:.3f} ms per step")
    E64 = batch.E.clone()
    dC64 = [d.clone() for d in batch.dC]
    t_fwd = timed("fp64", False, args.steps, 1)
    fk = batch.flops()
    # profiled eager step, systems one after another (no two kernels overlap,
    # so every launch's HIP events time that kernel alone): per-phase kernel
    # time summed over the 64 systems
    batch.set_profiling(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    batch.step("fp64", serial=True)
    torch.cuda.synchronize()
    t_serial = time.perf_counter() - t0
    phases = batch.phase_times()
    batch.set_profiling(False)
    # algorithmic work per phase of one step
    # the half-transform phase holds the forward pass, and the backward's
    # recompute of X for systems that do not keep Z = C^T L C (none here: Z fits)
    n_ht = [1.0 if z is not None else 2.0 for z in batch.Z]
    ht_f = sum(k * 2.0 * M * o * n * n for k, (n, M, o) in zip(n_ht, shapes))
    sy_f = sum(1.0 * M * o * n * (n + 1) for n, M, o in shapes)          # SYRK, lower triangle
    z_f = sum(2.0 * M * n * o * o for n, M, o in shapes)                 # Z = X C (forward)
    # W = C Z and the packed W C^T (fused), plus Z = X C where it is recomputed
    bw_f = sum(M * ((2.0 if z is not None else 4.0) * n * o * o + 1.0 * n * (n + 1) * o)
               for z, (n, M, o) in zip(batch.Z, shapes))
    ph = {
        "half_transform": ("k_ht<F64> (X = C^T L with the Coulomb V inside)", "k_ht", ht_f,
                           phases["half_transform"]),
        "exchange": ("k_syrk<F64> (K = sum X^T X)", "k_syrk", sy_f, phases["exchange"]),
        "backward": ("k_grad_rows (dE/dL: W = C Z and the packed W C^T in one launch)", "k_grad_rows",
                     bw_f, phases["backward"]),
        "zsave": ("k_gemm (Z = X C kept by the forward)", "k_gemm", z_f, phases["zsave"]),
    }
    name, (label, family, flops, (ms, calls)) = max(ph.items(), key=lambda kv: kv[1][3][0])
    achieved = flops / (ms * 1e-3) / 1e12
    # ladder: per-system split points for |dE| <= 1e-6 Eh, timed the same way
    progress("profiled eager step done")
    plans = batch.tune_ladder(tol=1e-6)
    progress("ladder tuned")
    t_lad = timed("ladder", True, args.steps, 1)
    Elad = batch.E.clone()
    dClad = [d.clone() for d in batch.dC]
    dE = torch.abs(Elad - E64)
    dCerr = max(torch.max(torch.abs(a - b)).item() for a, b in zip(dClad, dC64))
    m_tot = sum(M for _, M, _ in shapes)
    out = {
        "value": round(nsys / t64, 4),
        "ms_per_step": round(1e3 * t64, 3),
        "config": {
            "workload": "64 uneven systems (nbas U[100,600], nchol U[4,6] nbas, nocc U[0.1,0.2] nbas, "
                        "seed 5), forward J+K+E and backward dE/dC, dE/dL per system, one HIP graph replay",
            "systems": nsys, "nbas_total": sum(n for n, _, _ in shapes), "nchol_total": m_tot,
            "backward_from_saved_Z": sum(z is not None for z in batch.Z),
            "precision": "fp64", "parallelism": "1 GPU",
        },
        "roofline": {
            "bound": "mfma", "kernel": label, "phase": name,
            "achieved": round(achieved, 3), "peak": PEAK_TFLOPS["fp64"], "unit": "TFLOP/s",
            "frac": round(achieved / PEAK_TFLOPS["fp64"], 4),
            "kernel_ms_per_step": round(ms, 3), "launches_per_step": calls,
            "timing": "HIP events around every launch of one eager step of the same batch with the "
                      "systems run one after another (the graph replays the same launches, four lanes "
                      "at a time)",
            "traffic": batch_traffic(family, calls),
            "traffic_source": "profiles/batch64_fp64_fwd_bwd.md (FETCH_SIZE + WRITE_SIZE of the same step, "
                              "calibrated)",
            "phases_TFs": {k: round(v[2] / (v[3][0] * 1e-3) / 1e12, 2) for k, v in ph.items() if v[3][0] > 0},
        },
        "batch": {
            "fwd_bwd_replay_ms_fp64": round(1e3 * t64, 3),
            "fwd_replay_ms_fp64": round(1e3 * t_fwd, 3),
            "fwd_exchange_TFs_over_fwd_replay": round(fk / t_fwd / 1e12, 2),
            "fwd_exchange_gflop": round(fk / 1e9, 1),
            "phase_ms_per_step_eager_serial": {k: round(v[0], 3) for k, v in phases.items()},
            "eager_serial_step_ms": round(1e3 * t_serial, 3),
            "ladder": {
                "fwd_bwd_replay_ms": round(1e3 * t_lad, 3),
                "builds_per_s": round(nsys / t_lad, 4),
                "speedup_vs_fp64": round(t64 / t_lad, 3),
                "max_abs_dE_Eh": float(torch.max(dE).item()),
                "max_abs_dEdC_err": dCerr,
                "fp64_vector_fraction": round(sum(p.j_fp32 for p in plans) / m_tot, 4),
                # exchange work is n^2 o per vector: the FP64 head's cap with the
                # reduced-precision segments free
                "exchange_speedup_cap_fp64_head": round(ladder.speedup_bound(
                    plans, [n * n * o for n, _, o in shapes]), 3),
                "fp32_vector_fraction": round(sum(p.j_bf16 - p.j_fp32 for p in plans) / m_tot, 4),
                "tol_Eh": 1e-6,
            },
        },
    }
    batch.close()
    del batch
    torch.cuda.empty_cache()
    return out, shapes


def kernel_profile(dev, precision="fp64"):
    """The batch workload's kernels for a profiler: one serial eager step,
    preceded by a 0.5 s pause so that a kernel trace can cut it out.  With
    precision "ladder" the split points are tuned per system first."""
    import torch

    from openmolcas_amd.batch import MoleculeBatch, uneven_shapes

    batch = MoleculeBatch(uneven_shapes(64, seed=5), device=dev, seed=3)
    batch.autotune("fp64")
    extra = {}
    if precision == "ladder":
        plans = batch.tune_ladder(tol=1e-6)
        extra["fp64_vectors"] = int(sum(p.j_fp32 for p in plans))
        extra["fp32_vectors"] = int(sum(p.j_bf16 - p.j_fp32 for p in plans))
        extra["bf16x3_vectors"] = int(sum(M - p.j_bf16 for p, (n, M, o) in zip(plans, batch.shapes)))
    batch.capture(precision, True)
    batch.replay()
    torch.cuda.synchronize()
    time.sleep(0.5)
    batch.set_profiling(True)
    t0 = time.perf_counter()
    batch.step(precision, serial=True)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    phases = batch.phase_times()
    batch.close()
    label = "FP64" if precision == "fp64" else "precision-ladder"
    return {"workload": f"64-system batch, one {label} forward + backward eager step, systems one after another",
            "step_ms": round(1e3 * t, 3),
            "phase_ms": {k: round(v[0], 3) for k, v in phases.items()},
            "phase_launches": {k: v[1] for k, v in phases.items()}, **extra}


def single_build(args, dev, steps):
    """The synthetic single J/K build (nbas=512, nchol=2048, nocc=64 by
    default): FP64 fused build timed over `steps`, per-kernel phase times, the
    mixed modes and the ladder, the CPU oracle on the same build and the
    drop-in's parity against it."""
    import torch

    from openmolcas_amd.fock import JKContext
    from openmolcas_amd.synthetic import chol_vectors_torch, npair, orbitals_torch

    n, M, o = args.nbas, args.nchol, args.nocc
    npr = npair(n)

    L = chol_vectors_torch(n, M, args.seed, device=dev)
    C = orbitals_torch(n, o, args.seed + 1, device=dev)
    P = C @ C.T
    ii, jj = torch.tril_indices(n, n, device=dev)
    dlt = (P[ii, jj] * torch.where(ii == jj, 1.0, 2.0).to(torch.float64)).unsqueeze(0).contiguous()
    g = torch.Generator(device=dev)
    g.manual_seed(args.seed + 2)
    h = torch.randn((n, n), generator=g, dtype=torch.float64, device=dev)
    h = ((h + h.T) / (2.0 * n))[ii, jj].unsqueeze(0).contiguous()

    ctx = JKContext(n, M, device=dev)
    ctx.upload_packed(L)
    del L
    ctx.autotune([C], args.precision)  # tiles of the timed precision (HIP-event timed, median of 3)
    torch.cuda.synchronize()
    F = torch.empty((1, npr), dtype=torch.float64, device=dev)
    jsrc, fj, fk = [0], [1.0], [-0.5]

    # FP64: J and K in one pass over the vectors (V = L^T D inside the
    # half-transform), then the J = L V pass
    fused = args.precision == "fp64" and not args.two_pass

    def step():
        if fused:
            J, K = ctx.coulomb_exchange(dlt, [C])
        else:
            J = ctx.coulomb(dlt)
            K = ctx.exchange([C], precision=args.precision)
        F.copy_(h)
        ctx.fock_update(F, J, jsrc, fj, K, fk)
        return J, K

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ctx.set_profiling(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    phases = ctx.phase_times()
    ctx.set_profiling(False)
    elapsed = t1 - t0
    ms_step = 1000.0 * elapsed / steps

    # algorithmic work (per build)
    # the Coulomb passes of their own: V and J (two reads of L), or J alone when
    # V rides in the half-transform
    bytes_j = (1.0 if fused else 2.0) * M * npr * 8 + 2.0 * npr * 8 + 2.0 * M * 8
    flops_ht = 2.0 * M * o * n * n
    flops_syrk = 1.0 * M * o * n * (n + 1)
    peak = PEAK_TFLOPS[args.precision]

    def avg(name):
        ms, calls = phases[name]
        return ms / calls if calls else float("nan"), calls

    ht_ms, ht_calls = avg("half_transform")
    sy_ms, sy_calls = avg("exchange")
    jv_ms, jv_calls = avg("coulomb_v")
    if not jv_calls:
        jv_ms = 0.0
    jj_ms, _ = avg("coulomb_j")
    ht_per_launch = flops_ht / (ht_calls / steps)
    sy_per_launch = flops_syrk / (sy_calls / steps)
    ht_tf = ht_per_launch / (ht_ms * 1e-3) / 1e12
    sy_tf = sy_per_launch / (sy_ms * 1e-3) / 1e12
    j_gbs = bytes_j / ((jv_ms + jj_ms) * 1e-3) / 1e9
    k_ms = phases["half_transform"][0] / steps + phases["exchange"][0] / steps
    k_tf = (flops_ht + flops_syrk) / (k_ms * 1e-3) / 1e12
    # dominant kernel: the larger of the two exchange kernels by time
    # kernel names as the profiles list them (split-BF16 has kernels of its own)
    kht, ksy = {"fp64": ("k_ht<F64", "k_syrk<F64"), "fp32": ("k_ht<F32", "k_syrk<F32"),
                "bf16x3": ("k_ht_b3", "k_syrk_b3")}[args.precision]
    if ht_ms * ht_calls >= sy_ms * sy_calls:
        what = "half-transform X = C^T L" + (" with the Coulomb V = L^T D; MFMA flops only" if fused else "")
        dom = (f"{kht}> ({what})", ht_tf, kht + ("," if "<" in kht else "<"))
    else:
        dom = (f"{ksy}> (K = sum X^T X)", sy_tf, ksy + ("," if "<" in ksy else "<"))
    peaks = {name: round(ctx.measure_peak(kind), 2) for name, kind in
             (("mfma_f64_TFs", 0), ("mfma_f32_TFs", 1), ("mfma_bf16_TFs", 2), ("fma_f64_TFs", 3),
              ("copy_TBs", 4))}

    out = {
        "value": round(steps / elapsed, 4),
        "ms_per_step": round(ms_step, 4),
        "config": {
            "workload": f"synthetic single J/K build nbas={n} nchol={M} nocc={o}",
            "nbas": n, "nchol": M, "nocc": o, "precision": args.precision,
            "parallelism": "1 GPU",
        },
        "roofline": {
            "bound": "mfma",
            "kernel": dom[0],
            "achieved": round(dom[1], 3),
            "peak": peak,
            "unit": "TFLOP/s",
            "frac": round(dom[1] / peak, 4),
            "traffic": profiled_traffic(dom[2], n, M, o, args.precision),
            "peak_measured": peaks,
        },
        "kernels": {
            "coulomb_J_GBs": round(j_gbs, 1),
            "coulomb_J_frac_hbm": round(j_gbs / HBM_PEAK_GBS, 4),
            "coulomb_ms": round(jv_ms + jj_ms, 4),
            "coulomb_v_in_half_transform": fused,
            "half_transform_TFs": round(ht_tf, 2),
            "half_transform_ms": round(ht_ms, 4),
            "syrk_TFs": round(sy_tf, 2),
            "syrk_ms": round(sy_ms, 4),
            "exchange_TFs": round(k_tf, 2),
            "exchange_frac_peak": round(k_tf / peak, 4),
            "phase_ms_per_step": {k: round(v[0] / steps, 4) for k, v in phases.items()},
        },
    }
    if args.precision == "fp64" and not args.no_mixed:
        progress(f"single build n={n}: {ms_step:.4f} ms per step")
        out["mixed"] = mixed_vs_fp64(ctx, dlt, h, C, n, M, o, ms_step, args)
        progress(f"single build n={n}: mixed modes done")
    ctx.close()
    torch.cuda.empty_cache()
    if not args.no_cpu_baseline:
        out["cpu_baseline"], ref = cpu_baseline(n, M, o, args.seed, args.cpu_sample_vectors)
        out["parity"] = gpu_parity(ref, dev)
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: runs on one GPU only (asked for {args.gpus}); the J/K build would "
              "shard over Cholesky vectors with one all-reduce of the LT Fock matrix, which this "
              "library does not implement.", file=sys.stderr)
        sys.exit(2)
    import torch

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    line = {
        "metric": "J+K builds/sec + HBM GB/s (J), MFMA util (K): mixed vs FP64 at |dE|<=1e-6 Eh",
        "unit": "builds/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded N(0,1) symmetric vectors, norms 1..1e-6; QR orbitals; h = sym N(0,1)/nbas)",
    }
    if args.kernel_profile:
        print(json.dumps(kernel_profile(dev, args.profile_precision)), flush=True)
        return
    if args.workload == "batch64":
        head, shapes = batch_workload(args, dev)
        line.update(head)
        if not args.no_cpu_baseline:
            line["cpu_baseline"] = batch_cpu_baseline(shapes, seed=3)
            progress("cpu baseline done")
        if not args.no_secondary:
            single = single_build(args, dev, steps=20)
            progress("secondary n=512 build done")
            line["secondary"] = {"single_build_n512_m2048_o64": single, "build_n1200_m6000_o150": build1200(dev)}
            progress("secondary n=1200 build done")
    else:
        single = single_build(args, dev, steps=args.steps)
        line["dtype"] = "f64" if args.precision == "fp64" else args.precision
        line.update(single)
    if args.extra:
        line["extra"] = {"state_averaged_9": state_averaged(dev)}
    if args.copy_probe:
        src = torch.empty(1 << 27, dtype=torch.float64, device=dev)  # 1 GiB
        src.fill_(1.0)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
