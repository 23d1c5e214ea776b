"""Batched inference (many instances of uneven size in one model) and hipGraph replay.

Bars: each instance's lattice (V, offsets, neighbours) and Q are bit-identical to the
same image run alone through the plain path, and Q is within 1e-5 of the oracle; a
captured graph gives bit-identical Q on every replay, with eager work in between.
"""
import numpy as np
import pytest

from oracle import pyoracle as po
from pydensecrf_amd import capi
from pydensecrf_amd.batched import BatchedDenseCRF, BatchedDenseCRF2D
from tests.helpers import image_case

pytestmark = pytest.mark.gpu

SHAPES = [(37, 23), (64, 48), (31, 17), (80, 61), (16, 16), (45, 33)]  # odd sizes: pad lanes per instance


def _cases(L):
    return [image_case(w, h, L, seed=500 + i) for i, (w, h) in enumerate(SHAPES)]


def test_batched_2d_matches_single():
    L, n = 21, 5
    cases = _cases(L)
    b = BatchedDenseCRF2D(SHAPES, L)
    b.setUnaryEnergy([U for _, U, _ in cases])
    b.addPairwiseGaussian(sxy=3, compat=3)
    b.addPairwiseBilateral(sxy=30, srgb=13, rgbims=[img for img, _, _ in cases], compat=10)
    Qb = b.inference(n)
    for i, ((w, h), (img, U, _)) in enumerate(zip(SHAPES, cases)):
        s = capi.Model.create_2d(w, h, L)
        s.set_unary(U)
        s.add_gaussian(3.0, 3.0, 3.0)
        s.add_bilateral(30.0, 30.0, 13.0, 13.0, 13.0, img, 10.0)
        Qs = s.inference(n)
        assert np.array_equal(Qb[i].view(np.int32), np.ascontiguousarray(Qs).view(np.int32)), "instance %d" % i
        for k in range(2):
            lb = b.model.lattice(k, i, Ni=w * h)
            ls = s.lattice(k)
            assert lb["V"] == ls["V"]
            assert np.array_equal(lb["offset"], ls["offset"]) and np.array_equal(lb["nbr"], ls["nbr"])
        o = po.OracleCRF(W=w, H=h, L=L)
        o.set_unary(U)
        o.add_gaussian(3.0, 3.0, 3.0)
        o.add_bilateral(30.0, 30.0, 13.0, 13.0, 13.0, img, 10.0)
        assert np.abs(Qb[i] - o.inference(n)).max() <= 1e-5


def test_batched_generic_volumes():
    L, n = 4, 6
    rng = np.random.default_rng(9)
    dims = [(5, 8, 9), (7, 6, 5), (4, 9, 11)]
    feats_g, feats_b, Us, sizes = [], [], [], []
    for (Z, Y, X) in dims:
        zz, yy, xx = np.meshgrid(np.arange(Z), np.arange(Y), np.arange(X), indexing="ij")
        pos = np.stack([zz, yy, xx]).reshape(3, -1).astype(np.float32)
        inten = rng.uniform(0, 1, (1, Z * Y * X)).astype(np.float32)
        feats_g.append(pos / np.float32(3))
        feats_b.append(np.concatenate([pos / np.float32(4), inten / np.float32(0.1)]))
        Us.append(rng.standard_normal((L, Z * Y * X)).astype(np.float32))
        sizes.append(Z * Y * X)
    w = np.array([1, 2, 2, 3], np.float32)
    b = BatchedDenseCRF(sizes, L)
    b.setUnaryEnergy(Us)
    b.addPairwiseEnergy(feats_g, compat=w)
    b.addPairwiseEnergy(feats_b, compat=w)
    Qb = b.inference(n)
    for i in range(len(dims)):
        o = po.OracleCRF(N=sizes[i], L=L)
        o.set_unary(Us[i])
        o.add_pairwise(feats_g[i], w)
        o.add_pairwise(feats_b[i], w)
        assert np.abs(Qb[i] - o.inference(n)).max() <= 1e-5


def test_graph_replay_bitexact():
    import torch
    L, n = 21, 5
    cases = _cases(L)
    b = BatchedDenseCRF2D(SHAPES, L)
    b.setUnaryEnergy([U for _, U, _ in cases])
    b.addPairwiseGaussian(sxy=3, compat=3)
    b.addPairwiseBilateral(sxy=30, srgb=13, rgbims=[img for img, _, _ in cases], compat=10)
    ref = np.concatenate(b.inference(n), axis=1)
    Q = torch.zeros(b.N * L, device="cuda", dtype=torch.float32)
    b.capture(n, Q.data_ptr())
    for r in range(3):
        Q.fill_(-1.0)
        b.replay()
        torch.cuda.synchronize()
        got = Q.view(b.N, L).cpu().numpy().T
        assert np.array_equal(got.view(np.int32), ref.view(np.int32)), "replay %d" % r
        _ = b.inference(1)  # eager work between replays
