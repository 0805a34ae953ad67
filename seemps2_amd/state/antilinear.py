"""AntilinearForm <xi|psi> on the device (reference: src/seemps/state/antilinear.py:13-146)."""

from __future__ import annotations

from .. import kernels as K
from ..core import _begin_environment, _update_left_environment, _update_right_environment
from .mps import MPS


class AntilinearForm:
    """Environments of <bra|ket> around ``center`` and the one/two-site tensors
    whose contraction with the bra gives the scalar product."""

    def __init__(self, bra: MPS, ket: MPS, center: int = 0):
        if bra.size != ket.size:
            raise ValueError("bra and ket must have the same size")
        size = bra.size
        # a negative centre counts from the end, as the callers' start = -1 of a
        # right-to-left simplify means (antilinear.py:38-57 would build the
        # environments for a literal -1 and fail on the first two-site tensor)
        if center < 0:
            center += size
        if not 0 <= center < size:
            raise IndexError("AntilinearForm: centre out of range")
        rho = _begin_environment()
        R = [rho] * size
        for i in range(size - 1, center, -1):
            R[i - 1] = rho = _update_right_environment(bra[i], ket[i], rho)
        rho = _begin_environment()
        L = [rho] * size
        for i in range(0, center):
            L[i + 1] = rho = _update_left_environment(bra[i], ket[i], rho)
        self.bra, self.ket, self.size = bra, ket, size
        self.R, self.L, self.center = R, L, center

    def tensor1site(self):
        """L[c] . ket[c] . R[c] -> [bra_l, i, bra_r] (antilinear.py:59-66)."""
        c = self.center
        L, A, R = self.L[c], self.ket[c], self.R[c]
        a, d, b = A.shape
        AR = K.gemm(A.reshape(a * d, b), R)
        return K.gemm(L, AR.view(a, d * R.shape[1])).view(L.shape[0], d, R.shape[1])

    def tensor2site(self, direction: int, out=None, alpha=1.0, beta=0.0):
        """L[i] . ket[i] . ket[j] . R[j] -> [bra_l, i, j, bra_r] (antilinear.py:68-98).

        ``out``/``alpha``/``beta`` accumulate ``alpha * theta + beta * out`` in the
        last GEMM (used for sums of forms in simplify of an MPSSum)."""
        if direction > 0:
            i = self.center
            j = i + 1
        else:
            j = self.center
            i = j - 1
        L, A, B, R = self.L[i], self.ket[i], self.ket[j], self.R[j]
        a, d1, b = A.shape
        _, d2, c = B.shape
        xl, xr = L.shape[0], R.shape[1]
        LA = K.gemm(L, A.reshape(a, d1 * b))                        # [xl, (d1, b)]
        BR = K.gemm(B.reshape(b * d2, c), R)                        # [(b, d2), xr]
        if out is None:
            return K.gemm(LA.view(xl * d1, b), BR.view(b, d2 * xr), alpha=alpha).view(
                xl, d1, d2, xr)
        K.gemm(LA.view(xl * d1, b), BR.view(b, d2 * xr), alpha=alpha, beta=beta,
               C=out.view(xl * d1, d2 * xr))
        return out

    def update(self, direction: int) -> None:
        if direction > 0:
            self.update_right()
        else:
            self.update_left()

    def update_right(self) -> None:
        prev = self.center
        nxt = prev + 1
        if nxt >= self.size:
            raise RuntimeError("Cannot move the center beyond the last site")
        self.L[nxt] = _update_left_environment(self.bra[prev], self.ket[prev], self.L[prev])
        self.center = nxt

    def update_left(self) -> None:
        prev = self.center
        nxt = prev - 1
        if nxt < 0:
            raise RuntimeError("Cannot move the center before the first site")
        self.R[nxt] = _update_right_environment(self.bra[prev], self.ket[prev], self.R[prev])
        self.center = nxt
