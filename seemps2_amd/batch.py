"""Batched work over many MPS of uneven length and bond dimension.

``GroupedGemm`` is a fixed list of complex products C_p = op(A_p) op(B_p)
executed by one launch of the grouped GEMM kernel (``seemps_zgemm_grouped``):
the descriptors and the tile -> problem table are uploaded once, and the tile
shape (64 x 64 or 32 x 32) is chosen by ``TileTuner``, which times both on
the first list of each shape class and remembers the faster.

``BatchedOverlap`` computes <bra_k|ket_k> for many pairs of MPS at once.  Per
site index it runs two grouped GEMMs over every pair that still has that site
(the environment update rho'[n, k'] = sum conj(bra[l, j, n]) rho[l, i]
ket[i, j, k'] of the reference's scprod, src/seemps/cython/environments.pxi:
10-62, 87-100), so the whole batch costs two launches per site of the
longest state instead of two per site per pair.  The launch sequence is
captured once in a HIP graph (torch.cuda.CUDAGraph over the library's own
kernels) and replayed by ``run()``; the input tensors are referenced, not
copied, so updating them in place and replaying gives the new overlaps.
"""

from __future__ import annotations

import time

import numpy as np
import torch

from . import _lib
from . import kernels as K
from ._lib import check, lib, ptr, stream

OP_N, OP_T, OP_H, OP_C = 0, 1, 2, 3
TILE_64, TILE_32 = 0, 1
_DESC = np.dtype([("M", "i8"), ("N", "i8"), ("K", "i8"), ("A", "u8"), ("lda", "i8"),
                  ("B", "u8"), ("ldb", "i8"), ("C", "u8"), ("ldc", "i8"),
                  ("tiles_n", "i8"), ("tile0", "i8")])


class GroupedGemm:
    """One grouped launch for a fixed list of products.

    problems: list of (M, N, K, A, lda, B, ldb, C, ldc) with device pointers
    (ints) or tensors for A, B, C."""

    def __init__(self, opA: int, opB: int, problems, tile_cfg: int = TILE_64):
        self.opA, self.opB = opA, opB
        self.problems = [tuple(int(x.data_ptr()) if isinstance(x, torch.Tensor) else int(x)
                               for x in p) for p in problems]
        self.set_tile(tile_cfg)

    def set_tile(self, tile_cfg: int) -> None:
        self.tile_cfg = tile_cfg
        desc = np.zeros(len(self.problems), dtype=_DESC)
        prob_of_tile = []
        t0 = 0
        edge = 64 if tile_cfg == TILE_64 else 32
        for p, (M, N, Kd, A, lda, B, ldb, C, ldc) in enumerate(self.problems):
            if min(M, N) < 1 or Kd < 1:
                raise ValueError("GroupedGemm: empty product")
            nt = int(lib.seemps_grouped_tiles(tile_cfg, M, N))
            desc[p] = (M, N, Kd, A, lda, B, ldb, C, ldc, -(-N // edge), t0)
            prob_of_tile.append(np.full(nt, p, dtype=np.int32))
            t0 += nt
        self.ntiles = t0
        dev = _lib.device()
        self._desc = torch.from_numpy(desc.view(np.uint8)).to(dev)
        self._tiles = torch.from_numpy(np.concatenate(prob_of_tile) if prob_of_tile
                                       else np.zeros(0, np.int32)).to(dev)

    def launch(self) -> None:
        check(lib.seemps_zgemm_grouped(self.opA, self.opB, self.tile_cfg, len(self.problems),
                                       ptr(self._desc), ptr(self._tiles), self.ntiles, stream()))

    def signature(self):
        """Shape class for the tuner: op pair, problem count and median sizes,
        rounded to powers of two."""
        if not self.problems:
            return (self.opA, self.opB, 0)
        sizes = np.array([p[:3] for p in self.problems], dtype=np.float64)
        med = np.exp2(np.round(np.log2(np.median(sizes, axis=0))))
        return (self.opA, self.opB, int(np.exp2(np.round(np.log2(len(self.problems))))),
                *map(int, med))


class TileTuner:
    """Picks the grouped GEMM tile shape per shape class by timing both
    configurations (HIP events on the current stream) the first time a class
    is seen."""

    def __init__(self, reps: int = 3):
        self.reps = reps
        self.choices: dict = {}

    def _time(self, gg: GroupedGemm) -> float:
        start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        gg.launch()
        start.record()
        for _ in range(self.reps):
            gg.launch()
        stop.record()
        stop.synchronize()
        return start.elapsed_time(stop) / self.reps

    def tune(self, gg: GroupedGemm) -> int:
        sig = gg.signature()
        cfg = self.choices.get(sig)
        if cfg is None:
            times = {}
            for c in (TILE_64, TILE_32):
                gg.set_tile(c)
                times[c] = self._time(gg)
            cfg = min(times, key=times.get)
            self.choices[sig] = cfg
        if gg.tile_cfg != cfg:
            gg.set_tile(cfg)
        return cfg


DEFAULT_TUNER = TileTuner()


class BatchedOverlap:
    """<bra_k|ket_k> for many MPS pairs of uneven length and bond dimension,
    as one replayable HIP graph of grouped environment updates."""

    def __init__(self, bras, kets, tuner: TileTuner | None = DEFAULT_TUNER,
                 graph: bool = True):
        if len(bras) != len(kets) or not bras:
            raise ValueError("BatchedOverlap: need as many bras as kets (at least one)")
        self.pairs = []
        for b, k in zip(bras, kets):
            bt, kt = [K.as_dev(A) for A in b], [K.as_dev(A) for A in k]
            if len(bt) != len(kt):
                raise ValueError("BatchedOverlap: bra and ket lengths differ")
            for A, B in zip(bt, kt):
                if A.shape[1] != B.shape[1]:
                    raise ValueError("BatchedOverlap: physical dimensions differ")
            if bt[0].shape[0] != 1 or kt[0].shape[0] != 1 or bt[-1].shape[2] != 1 or \
                    kt[-1].shape[2] != 1:
                raise ValueError("BatchedOverlap: open boundary bonds must be 1")
            self.pairs.append((bt, kt))
        n = len(self.pairs)
        self.out = K.empty(n)
        # per pair: rho (chi_bra x chi_ket) and T (chi_bra x d chi_ket) buffers
        self._rho, self._T = [], []
        for bt, kt in self.pairs:
            cb = max(A.shape[2] for A in bt)
            ck = max(A.shape[2] for A in kt)
            dk = max(A.shape[1] * A.shape[2] for A in kt)
            self._rho.append(K.empty(max(cb * ck, 1)))
            self._T.append(K.empty(max(cb * dk, 1)))
        self.stages = []
        longest = max(len(bt) for bt, _ in self.pairs)
        for s in range(longest):
            first, second = [], []
            for p, (bt, kt) in enumerate(self.pairs):
                if s >= len(bt):
                    continue
                A, B = bt[s], kt[s]
                a, d, b = A.shape           # bra site: chi_bra_l, d, chi_bra_r
                i, _, k = B.shape           # ket site: chi_ket_l, d, chi_ket_r
                if s == 0:
                    T = B                   # rho = [[1]]: T = ket[0]
                else:
                    T = self._T[p]
                    # T[l, (j, k')] = rho[l, i] ket[i, (j, k')]
                    first.append((a, d * k, i, self._rho[p], i, B, d * k, T, d * k))
                last = s == len(bt) - 1
                C = self.out.data_ptr() + 16 * p if last else self._rho[p].data_ptr()
                # rho'[n, k'] = sum_(l,j) conj(bra[(l,j), n]) T[(l,j), k']
                second.append((b, k, a * d, A, b, T, k, C, k))
            st = []
            if first:
                st.append(GroupedGemm(OP_N, OP_N, first))
            st.append(GroupedGemm(OP_H, OP_N, second))
            self.stages.append(st)
        if tuner is not None:
            # every stage of the chain reads what the previous one wrote: time in order
            for st in self.stages:
                for gg in st:
                    tuner.tune(gg)
                    gg.launch()
        self.graph = None
        if graph:
            self._launch_all()  # warm-up outside capture
            torch.cuda.synchronize()
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                self._launch_all()
            self.graph = g

    def _launch_all(self) -> None:
        for st in self.stages:
            for gg in st:
                gg.launch()

    def launches(self) -> int:
        return sum(len(st) for st in self.stages)

    def run_async(self) -> torch.Tensor:
        """Enqueue the overlaps on the current stream; returns the device vector."""
        if self.graph is not None:
            self.graph.replay()
        else:
            self._launch_all()
        return self.out

    def run(self) -> np.ndarray:
        return self.run_async().cpu().numpy()


def batched_scprod(bras, kets, tuner: TileTuner | None = DEFAULT_TUNER) -> np.ndarray:
    """<bra_k|ket_k> for every pair (one-off: no graph)."""
    return BatchedOverlap(bras, kets, tuner=tuner, graph=False).run()


__all__ = ["BatchedOverlap", "GroupedGemm", "TileTuner", "batched_scprod"]
