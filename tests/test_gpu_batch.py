"""Grouped GEMM and batched overlaps of many uneven MPS (device)."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def S():
    import seemps2_amd

    return seemps2_amd


def crandn(rng, *shape):
    return rng.normal(size=shape) + 1j * rng.normal(size=shape)


@pytest.mark.parametrize("tile", [0, 1])
def test_grouped_gemm_uneven(S, rng, tile):
    from seemps2_amd import batch, kernels as K

    OPS = {0: lambda x: x, 1: lambda x: x.T, 2: lambda x: x.conj().T, 3: lambda x: x.conj()}
    for opA, opB in [(0, 0), (2, 0), (0, 1), (1, 0), (3, 0), (0, 2)]:
        probs, refs, outs, keep = [], [], [], []
        for _ in range(7):
            M, N, Kd = (int(x) for x in rng.integers(1, 150, size=3))
            A = crandn(rng, *((Kd, M) if opA in (1, 2) else (M, Kd)))
            B = crandn(rng, *((N, Kd) if opB in (1, 2) else (Kd, N)))
            Ad, Bd, Cd = K.as_dev(A), K.as_dev(B), K.empty(M, N)
            keep += [Ad, Bd]
            probs.append((M, N, Kd, Ad, A.shape[1], Bd, B.shape[1], Cd, N))
            refs.append(OPS[opA](A) @ OPS[opB](B))
            outs.append(Cd)
        gg = batch.GroupedGemm(opA, opB, probs, tile)
        gg.launch()
        for Cd, ref in zip(outs, refs):
            assert np.linalg.norm(K.to_numpy(Cd) - ref) <= 1e-14 * np.linalg.norm(ref) * 10


def _pairs(S, rng, count, same=False):
    from oracle import builders as ob

    bras, kets = [], []
    for _ in range(count):
        N = int(rng.integers(2, 9))
        a = ob.random_mps([2] * N, D=int(rng.integers(1, 20)), complex=True, rng=rng)
        b = a if same else ob.random_mps([2] * N, D=int(rng.integers(1, 20)), complex=True,
                                         rng=rng)
        bras.append(a)
        kets.append(b)
    return bras, kets


def test_batched_overlap_matches_scprod(S, rng):
    from seemps2_amd import batch

    bras, kets = _pairs(S, rng, 23)
    db = [S.MPS(a.tensors) for a in bras]
    dk = [S.MPS(b.tensors) for b in kets]
    ref = np.array([np.vdot(a.to_vector(), b.to_vector()) for a, b in zip(bras, kets)])
    for graph in (False, True):
        bo = batch.BatchedOverlap(db, dk, graph=graph)
        out = bo.run()
        assert np.abs(out - ref).max() <= 1e-12 * np.abs(ref).max()
        assert bo.launches() <= 2 * max(a.size for a in bras)
    # single-state scprod agrees too
    assert abs(complex(S.scprod(db[0], dk[0])) - ref[0]) <= 1e-12 * abs(ref[0]) + 1e-300
    # both tile shapes give the same overlaps
    t64 = batch.BatchedOverlap(db, dk, tuner=None, graph=False).run()
    bo32 = batch.BatchedOverlap(db, dk, tuner=None, graph=False)
    for st in bo32.stages:
        for gg in st:
            gg.set_tile(batch.TILE_32)
    assert np.abs(bo32.run() - t64).max() <= 1e-13 * np.abs(t64).max()


def test_batched_overlap_graph_replays(S, rng):
    """The captured graph reads the live tensors: updating a ket in place and
    replaying (several times, with other device work in between) gives the
    new overlaps every time."""
    from seemps2_amd import batch, kernels as K

    bras, kets = _pairs(S, rng, 9, same=True)
    db = [S.MPS(a.tensors) for a in bras]
    dk = [S.MPS(b.tensors) for b in kets]
    bo = batch.BatchedOverlap(db, dk, graph=True)
    norms = np.array([np.vdot(a.to_vector(), a.to_vector()).real for a in bras])
    for rep in range(4):
        factor = 2.0 ** rep
        out = bo.run()
        assert np.abs(out - factor * norms).max() <= 1e-12 * factor * norms.max()
        # eager work between replays, then scale every ket's first tensor by 2 in place
        _ = S.scprod(db[0], dk[0])
        for k in dk:
            K.axpby(0.0, None, 2.0, k[0])
        torch.cuda.synchronize()


def test_tuner_remembers_choice(S, rng):
    from seemps2_amd import batch

    bras, kets = _pairs(S, rng, 5)
    tuner = batch.TileTuner()
    batch.BatchedOverlap([S.MPS(a.tensors) for a in bras], [S.MPS(b.tensors) for b in kets],
                         tuner=tuner, graph=False)
    assert tuner.choices and set(tuner.choices.values()) <= {batch.TILE_64, batch.TILE_32}
