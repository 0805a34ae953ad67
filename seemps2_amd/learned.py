"""Learned-rank truncation for the two-site splits.

A small MLP (32 -> 64 -> 64 -> 1, ReLU, FP32) predicts the fraction of the
singular values a bond keeps from a cheap spectral sketch of the two-site
matrix (``seemps_spectral_sketch``: 32 random sketch columns, the log-spectrum
of their Gram matrix).  The split then uses a fixed-rank randomized SVD
(``seemps_svd_split_sketch``) with that many columns plus a safety margin; the
full Jacobi SVD is the fallback whenever the sketch misses more weight than the
Strategy may discard, so the kept rank is never decided on a sketch that is
too small.  Training data are the spectra of the plain path's own full SVDs
(:class:`SpectrumLog`), and training runs on the device through the HIP
forward, backward and Adam kernels (``seemps_mlp_*``).

When the split comes with the bond's current basis (the canonical neighbour
that a DMRG sweep is about to replace), the sketch is that basis plus
p = max(oversample, l - k) expansion directions (``seemps_svd_split_basis``):
the old basis already spans almost all of the new two-site matrix, so the
Jacobi SVD of the sketched matrix needs neither a QR preconditioner nor many
sweeps, and the predicted rank decides how far the bond may grow beyond it.

Nothing here changes the reference-compatible default path: the learned
splits are used only inside ``with learned_truncation(...)``.
"""

from __future__ import annotations

import ctypes
import math
from contextlib import contextmanager
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from . import kernels as K
from ._lib import check, lib, ptr, stream

F_IN = 32
HIDDEN = 64
#: Largest number of expansion directions of a basis-sketched split.
MAX_EXPANSION = 32


def spectral_features(AA: torch.Tensor, m: int, n: int, seed: int = 0) -> torch.Tensor:
    """The 32 FP32 predictor inputs of an m x n complex matrix (device tensor)."""
    AA = K.as_dev(AA)
    out = torch.empty(F_IN, dtype=torch.float32, device=AA.device)
    wp, wn = _lib.workspace.get(lib.seemps_spectral_sketch_workspace(m, n))
    check(lib.seemps_spectral_sketch(m, n, ptr(AA), seed & 0xFFFFFFFFFFFFFFFF, ptr(out), wp, wn,
                                     stream()))
    return out


class RankPredictor:
    """The MLP: parameters live on the device; all arithmetic runs in the HIP
    kernels.  The output is the predicted kept fraction D / min(m, n)."""

    def __init__(self, seed: int = 0, params: np.ndarray | None = None):
        n = lib.seemps_mlp_param_count(F_IN, HIDDEN)
        if params is None:
            rng = np.random.default_rng(seed)
            W1 = rng.normal(scale=math.sqrt(2.0 / F_IN), size=(HIDDEN, F_IN))
            W2 = rng.normal(scale=math.sqrt(2.0 / HIDDEN), size=(HIDDEN, HIDDEN))
            w3 = rng.normal(scale=math.sqrt(1.0 / HIDDEN), size=HIDDEN)
            params = np.concatenate([W1.ravel(), np.zeros(HIDDEN), W2.ravel(), np.zeros(HIDDEN),
                                     w3, np.zeros(1)])
        params = np.asarray(params, dtype=np.float32)
        if params.size != n:
            raise ValueError(f"RankPredictor: expected {n} parameters, got {params.size}")
        self.params = torch.tensor(params, device=_lib.device())

    def state_dict(self) -> dict:
        return {"params": self.params.cpu().numpy()}

    @classmethod
    def from_state_dict(cls, state: dict) -> "RankPredictor":
        return cls(params=state["params"])

    def _forward(self, X: torch.Tensor):
        B = X.shape[0]
        h1 = torch.empty(B, HIDDEN, dtype=torch.float32, device=X.device)
        h2 = torch.empty_like(h1)
        y = torch.empty(B, dtype=torch.float32, device=X.device)
        check(lib.seemps_mlp_forward(B, F_IN, HIDDEN, ptr(self.params), ptr(X), ptr(h1), ptr(h2),
                                     ptr(y), stream()))
        return h1, h2, y

    def predict(self, X: torch.Tensor) -> torch.Tensor:
        """Predicted kept fraction for each row of X (B x 32, FP32)."""
        X = X.to(device=self.params.device, dtype=torch.float32).reshape(-1, F_IN).contiguous()
        return self._forward(X)[2]

    def loss_and_grad(self, X: torch.Tensor, t: torch.Tensor):
        """Mean squared error and its gradient (device kernels)."""
        X = X.to(torch.float32).contiguous()
        t = t.to(device=X.device, dtype=torch.float32).contiguous()
        B = X.shape[0]
        h1, h2, y = self._forward(X)
        grads = torch.empty_like(self.params)
        loss = torch.empty(1, dtype=torch.float32, device=X.device)
        scratch = torch.empty(lib.seemps_mlp_scratch_count(B, HIDDEN), dtype=torch.float32,
                              device=X.device)
        check(lib.seemps_mlp_backward(B, F_IN, HIDDEN, ptr(self.params), ptr(X), ptr(h1),
                                      ptr(h2), ptr(y), ptr(t), ptr(grads), ptr(loss),
                                      ptr(scratch), stream()))
        return loss, grads

    def train(self, X: torch.Tensor, t: torch.Tensor, epochs: int = 400, lr: float = 3e-3,
              batch: int | None = None, seed: int = 0) -> list[float]:
        """Adam on the mean squared error; returns the loss per epoch."""
        X = X.to(device=self.params.device, dtype=torch.float32).contiguous()
        t = t.to(device=self.params.device, dtype=torch.float32).contiguous()
        m = torch.zeros_like(self.params)
        v = torch.zeros_like(self.params)
        N = X.shape[0]
        batch = N if batch is None else min(batch, N)
        g = torch.Generator(device="cpu").manual_seed(seed)
        losses = []
        step = 0
        for _ in range(epochs):
            order = torch.randperm(N, generator=g).to(X.device) if batch < N else None
            total = 0.0
            for s0 in range(0, N, batch):
                idx = order[s0:s0 + batch] if order is not None else None
                xb = X[idx].contiguous() if idx is not None else X
                tb = t[idx].contiguous() if idx is not None else t
                loss, grads = self.loss_and_grad(xb, tb)
                step += 1
                check(lib.seemps_adam_step(self.params.numel(), ptr(self.params), ptr(grads),
                                           ptr(m), ptr(v), lr, 0.9, 0.999, 1e-8, step, stream()))
                total += float(loss.item()) * xb.shape[0]
            losses.append(total / N)
        return losses


@dataclass
class SpectrumLog:
    """Training pairs (sketch features, kept fraction) from full-SVD splits."""

    features: list = field(default_factory=list)
    targets: list = field(default_factory=list)
    min_size: int = 32

    def add(self, feats: torch.Tensor, D: int, r: int) -> None:
        self.features.append(feats.clone())
        self.targets.append(D / r)

    def tensors(self):
        X = torch.stack(self.features) if self.features else torch.empty(0, F_IN)
        return X, torch.tensor(self.targets, dtype=torch.float32, device=X.device)


@dataclass
class LearnedTruncation:
    """Decides the sketch width of each two-site split from the predictor.

    l = ceil(margin * predicted kept rank) + oversample, capped by
    max_bond + oversample; splits smaller than ``min_size`` use the full SVD."""

    predictor: RankPredictor
    oversample: int = 16
    margin: float = 1.25
    power_iters: int = 2
    min_size: int = 64
    seed: int = 1234
    use_basis: bool = True
    #: after this many sketched splits in a row fell back to the full SVD (a
    #: spectrum the sketches cannot certify, such as the flat spectra of
    #: random circuits), the next ``backoff`` splits go to the full SVD
    #: directly; the pause doubles after every streak, up to 64 times itself
    fallback_streak: int = 8
    backoff: int = 8
    calls: int = 0
    fallbacks: int = 0
    full_splits: int = 0
    basis_splits: int = 0
    sketch_columns: int = 0
    _streak: int = 0
    _skip: int = 0
    _pause: int = 0

    def _record(self, fb: bool) -> None:
        self.fallbacks += int(fb)
        if not fb:
            self._streak = 0
            self._pause = 0
            return
        self._streak += 1
        if self._streak >= self.fallback_streak:
            self._pause = min(max(self.backoff, 2 * self._pause), 64 * self.backoff)
            self._skip = self._pause
            self._streak = 0

    def predicted_fraction(self, feats: torch.Tensor) -> float:
        frac = float(self.predictor.predict(feats.reshape(1, -1))[0].item())
        return min(max(frac, 0.0), 1.0)

    def sketch_width(self, feats: torch.Tensor, m: int, n: int, max_bond: int,
                     frac: float | None = None) -> int:
        r = min(m, n)
        if frac is None:
            frac = self.predicted_fraction(feats)
        l = int(math.ceil(self.margin * frac * r)) + self.oversample
        return min(l, max_bond + self.oversample, r)

    def split(self, AA2: torch.Tensor, m: int, n: int, mode: int, strategy, basis=None):
        """(A, B, D, err^2) of the two-site split, by sketch when it pays.

        ``basis``: the canonical neighbour the split replaces (mode 0: the
        right site, k x d x b, right-canonical; mode 1: the left site,
        a x d x k, left-canonical), or None."""
        r = min(m, n)
        self.calls += 1
        if r < self.min_size or self._skip > 0:
            if r >= self.min_size:
                self._skip -= 1
            self.full_splits += 1
            return K.svd_split(AA2, m, n, mode, strategy.method, strategy.tolerance,
                               strategy.max_bond_dimension)[:4]
        k = None
        if basis is not None and self.use_basis:
            k = basis.shape[0] if mode == 0 else basis.shape[-1]
            if strategy.max_bond_dimension <= k < r:
                # a bond at the cap cannot grow: the basis alone is the sketch
                # (the truncation certifies it, or the full SVD decides), and
                # no width needs predicting
                return self._basis_split(AA2, m, n, mode, basis, k, 0, strategy)
        feats = spectral_features(AA2, m, n, seed=self.seed + self.calls)
        frac = self.predicted_fraction(feats)
        l = self.sketch_width(feats, m, n, strategy.max_bond_dimension, frac)
        need = min(int(math.ceil(frac * r)), strategy.max_bond_dimension)
        if k is not None and need - k <= MAX_EXPANSION:
            # at most MAX_EXPANSION new directions per split: the expansion
            # block is then rotated to orthogonal images in one workgroup and
            # the sketched SVD stays the fast rotation iteration.  A bond
            # predicted to grow by more skips the basis (its split would fall
            # back); one that grows more than predicted is caught by the
            # truncation (the weight outside the sketch exceeds the budget)
            # and split in full
            # as many directions as the predicted growth of the bond (the
            # unmargined prediction less the basis), at least the oversample:
            # a bond that holds its size needs only the oversample, and every
            # extra direction lengthens the sketched SVD and the expansion
            # rotation
            p = min(max(self.oversample, need - k), MAX_EXPANSION)
            if k + p < r:
                return self._basis_split(AA2, m, n, mode, basis, k, p, strategy)
        if l >= r:
            self.full_splits += 1
            return K.svd_split(AA2, m, n, mode, strategy.method, strategy.tolerance,
                               strategy.max_bond_dimension)[:4]
        A, B, D, err2, _, fb = K.svd_split_sketch(AA2, m, n, mode, strategy.method,
                                                  strategy.tolerance,
                                                  strategy.max_bond_dimension, l,
                                                  power_iters=self.power_iters,
                                                  seed=self.seed + 7919 * self.calls)
        self.sketch_columns += l
        self._record(fb)
        return A, B, D, err2

    def _basis_split(self, AA2, m, n, mode, basis, k, p, strategy):
        bm = basis.reshape(k, n) if mode == 0 else basis.reshape(m, k)
        A, B, D, err2, _, fb = K.svd_split_basis(
            AA2, m, n, mode, bm, k, p, mode, strategy.method, strategy.tolerance,
            strategy.max_bond_dimension, seed=self.seed + 7919 * self.calls)
        self.basis_splits += 1
        self.sketch_columns += k + p
        self._record(fb)
        return A, B, D, err2


_ACTIVE: LearnedTruncation | None = None
_RECORDING: SpectrumLog | None = None
#: Two-site splits performed so far, in every form (for per-update rates).
SPLITS = 0


@contextmanager
def learned_truncation(lt: LearnedTruncation):
    """Two-site splits inside the block use the learned sketch path."""
    global _ACTIVE
    prev, _ACTIVE = _ACTIVE, lt
    try:
        yield lt
    finally:
        _ACTIVE = prev


@contextmanager
def recording(log: SpectrumLog):
    """Two-site splits inside the block log (features, kept fraction) pairs."""
    global _RECORDING
    prev, _RECORDING = _RECORDING, log
    try:
        yield log
    finally:
        _RECORDING = prev


def two_site_split(AA2: torch.Tensor, m: int, n: int, mode: int, strategy, basis=None):
    """The split used by the two-site updates: plain, learned or recorded.
    ``basis`` is the optional canonical neighbour (see LearnedTruncation.split);
    the plain split ignores it."""
    global SPLITS
    SPLITS += 1
    if _ACTIVE is not None:
        return _ACTIVE.split(AA2, m, n, mode, strategy, basis)
    A, B, D, err2, _ = K.svd_split(AA2, m, n, mode, strategy.method, strategy.tolerance,
                                   strategy.max_bond_dimension)
    if _RECORDING is not None and min(m, n) >= _RECORDING.min_size:
        _RECORDING.add(spectral_features(AA2, m, n, seed=len(_RECORDING.targets)), D, min(m, n))
    return A, B, D, err2
