"""Batched J/K builds for many independent systems, captured as one HIP graph.

Each system (molecule) has its own context -- its Cholesky vectors live in that
context's HBM store, written in place.  The systems are dealt to ``lanes``
streams (largest first, each to the lane with the least work so far); the
contexts of one lane share one set of device workspaces and that lane's stream
(mjk_create_shared), so the lanes run concurrently and small systems fill the
GPU together.  ``step()`` runs, for every system, the closed-shell energy build
(density, Coulomb, exchange, Fock and energy steps) followed, unless
``backward=False``, by the hand-written backward (dE/dC = 4FC and the packed
dE/dL); in FP64 the Coulomb V pass runs inside the half-transform
(mjk_coulomb_exchange).  When dE/dL is wanted and a system's nchol nocc^2
doubles fit ``z_budget`` bytes, its forward also keeps Z^J = C^T L^J C
(mjk_coulomb_exchange_z / mjk_exchange_z) and the backward forms dE/dL from Z
without a second half-transform; otherwise the backward recomputes X^J.  ``tune_ladder`` places each system's FP64 / FP32 /
split-BF16 split points for an energy tolerance (openmolcas_amd.ladder.tune,
verified per system) and precision "ladder" runs them.  ``capture()``
records one step -- the fork to the lane streams and the join back included --
into a torch CUDAGraph (a hipGraph on ROCm) after a warm-up step has sized
every workspace, so a replay is a single graph launch with no allocation, no
host synchronisation and no memset nodes.
"""
from __future__ import annotations

import numpy as np
import torch

from .fock import JKContext
from .synthetic import chol_vectors_torch, npair, orbitals_torch


class MoleculeBatch:
    def __init__(self, shapes, device=None, seed: int = 0, with_grad_l: bool = True,
                 lanes: int = 4, z_budget: int = 16 << 30):
        """shapes: list of (nbas, nchol, nocc)."""
        self.device = torch.device(device) if device is not None else torch.device("cuda", 0)
        self.stream = torch.cuda.Stream(device=self.device)
        self.shapes = [tuple(int(v) for v in s) for s in shapes]
        nl = max(1, min(int(lanes), len(self.shapes)))
        self.lane_streams = [torch.cuda.Stream(device=self.device) for _ in range(nl)]
        # deal systems to lanes: largest exchange work first, to the least loaded lane
        work = [M * o * n * n for n, M, o in self.shapes]
        load = [0.0] * nl
        self.lane_of = [0] * len(self.shapes)
        for i in sorted(range(len(self.shapes)), key=lambda i: -work[i]):
            k = min(range(nl), key=lambda k: load[k])
            self.lane_of[i] = k
            load[k] += work[i]
        self.ctx = [None] * len(self.shapes)
        self.L, self.C, self.h = [None] * len(shapes), [None] * len(shapes), [None] * len(shapes)
        self.E = torch.zeros(len(shapes), dtype=torch.float64, device=self.device)
        self.dC, self.dL, self.Z = [None] * len(shapes), [None] * len(shapes), [None] * len(shapes)
        first = [None] * nl
        for i, (n, M, o) in enumerate(self.shapes):
            k = self.lane_of[i]
            with torch.cuda.stream(self.lane_streams[k]):
                c = JKContext(n, M, device=self.device, share=first[k])
                if first[k] is None:
                    first[k] = c
                    c.bind_stream(self.lane_streams[k])
                self.ctx[i] = c
                L = c.device_vectors()
                chol_vectors_torch(n, M, seed + 7 * i, device=self.device, out=L)
                self.L[i] = L
                self.C[i] = orbitals_torch(n, o, seed + 7 * i + 1, device=self.device)
                g = torch.Generator(device=self.device)
                g.manual_seed(seed + 7 * i + 2)
                hs = torch.randn((n, n), generator=g, dtype=torch.float64, device=self.device)
                ii, jj = torch.tril_indices(n, n, device=self.device)
                self.h[i] = ((hs + hs.T) / (2.0 * n))[ii, jj].reshape(1, -1).contiguous()
                self.dC[i] = torch.zeros((n, o), dtype=torch.float64, device=self.device)
                self.dL[i] = (torch.zeros((M, npair(n)), dtype=torch.float64, device=self.device)
                              if with_grad_l else None)
                if with_grad_l and 8 * M * o * o <= z_budget:
                    self.Z[i] = torch.zeros((M, o, o), dtype=torch.float64, device=self.device)
        for st in self.lane_streams:
            st.synchronize()
        self.graph = None
        self.ladders = []   # per-system ladder plans (tune_ladder)

    def autotune(self, precision="fp64"):
        for i, (c, C) in enumerate(zip(self.ctx, self.C)):
            with torch.cuda.stream(self.lane_streams[self.lane_of[i]]):
                c.autotune([C], precision)
        for st in self.lane_streams:
            st.synchronize()

    def tune_ladder(self, tol: float = 1e-6, accept: float = 0.5):
        """Per system, the precision ladder whose measured |dE| against the
        FP64 build is at most accept * tol (E = h.D + 1/2 (J - K[C C^T]).D with
        D = fold(2 C C^T), so dE = 1/2 dK.D).  Sets each context's split points
        for precision "ladder"; returns the plans."""
        from . import ladder

        self.ladders = []
        for i, (c, C) in enumerate(zip(self.ctx, self.C)):
            with torch.cuda.stream(self.lane_streams[self.lane_of[i]]):
                dlt = c.density(C, 2.0)
                plan = ladder.tune(c, [C], dlt, factor=0.5, tol=tol, accept=accept)
                c.set_ladder(plan.j_fp32, plan.j_bf16)
                self.ladders.append(plan)
        for st in self.lane_streams:
            st.synchronize()
        return self.ladders

    def _system(self, i, precision, backward=True):
        c, C, h = self.ctx[i], self.C[i], self.h[i]
        Z = self.Z[i] if backward else None          # Z^J = C^T L^J C for dE/dL
        dlt = c.density(C, 2.0)                      # fold(P), P = 2 C C^T
        if precision == "fp64":                      # V pass inside the half-transform
            J, K, V = c.coulomb_exchange(dlt, [C], want_v=True, z=Z)
        else:
            J, V = c.coulomb(dlt, want_v=True)
            K = c.exchange([C], precision=precision, z=Z)  # K[C C^T]
        G = torch.zeros_like(J)
        c.fock_update(G, J, [0], [1.0], K, [-1.0])   # G = J - 1/2 K[P]
        self.E[i:i + 1].copy_(c.energy(h, dlt, G))
        if not backward:
            return
        F = h + G
        self.dC[i].copy_(c.grad_orbitals(C, F, 4.0))
        if self.dL[i] is not None:
            # past the ladder's FP64 head the products C Z^J C^T are formed in FP32
            tail = self.ladders[i].j_fp32 if precision == "ladder" and Z is not None and self.ladders else None
            c.grad_vectors(C, V[:, 0].contiguous(), 1.0, -2.0, out=self.dL[i], z=Z, fp32_from=tail)

    def step(self, precision="fp64", backward=True, serial=False):
        """Forward (+ backward) of every system: forks from the batch stream to
        the lane streams and joins back (all queued, nothing synchronised).
        serial=True makes every system wait for the one queued before it, so
        no two systems' kernels overlap (per-kernel timing)."""
        with torch.cuda.stream(self.stream):
            for st in self.lane_streams:
                st.wait_stream(self.stream)
            prev = None
            for i in range(len(self.shapes)):
                st = self.lane_streams[self.lane_of[i]]
                if serial and prev is not None and prev is not st:
                    st.wait_stream(prev)
                with torch.cuda.stream(st):
                    self._system(i, precision, backward)
                prev = st
            for st in self.lane_streams:
                self.stream.wait_stream(st)

    def capture(self, precision="fp64", backward=True):
        """Warm-up step (sizes the shared workspaces), then one captured step."""
        self.step(precision, backward)
        self.stream.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph, stream=self.stream):
            self.step(precision, backward)
        self.stream.synchronize()
        return self.graph

    def replay(self):
        self.graph.replay()

    def flops(self):
        """Algorithmic exchange flop of one forward pass: sum M o (2n^2 + n(n+1))."""
        return float(sum(M * o * (2 * n * n + n * (n + 1)) for n, M, o in self.shapes))

    def set_profiling(self, on: bool):
        for c in self.ctx:
            c.set_profiling(on)

    def phase_times(self):
        """Per-phase milliseconds and launches summed over every system."""
        tot = {}
        for c in self.ctx:
            for k, (ms, calls) in c.phase_times().items():
                a, b = tot.get(k, (0.0, 0))
                tot[k] = (a + ms, b + calls)
        return tot

    def close(self):
        for c in reversed(self.ctx):
            c.close()


def uneven_shapes(count: int, seed: int):
    """count systems, nbas ~ U[100,600], nchol = round(k nbas) with k ~ U[4,6],
    nocc = round(U[0.1,0.2] nbas)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n = int(rng.integers(100, 601))
        M = int(round(rng.uniform(4.0, 6.0) * n))
        o = max(1, int(round(rng.uniform(0.1, 0.2) * n)))
        out.append((n, M, o))
    return out
