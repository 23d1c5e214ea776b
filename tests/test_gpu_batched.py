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


def test_batched_instance_unaries_partial_and_reset():
    """Unaries given per instance: an instance never given one computes with a zero
    unary (as a plain model given zeros), a later full unary replaces them all, and a
    large instance (its label rows copied by several threads) is bit-identical to its
    image run alone."""
    L, n = 21, 3
    shapes = [(37, 23), (640, 480), (31, 17)]
    cases = [image_case(w, h, L, seed=700 + i) for i, (w, h) in enumerate(shapes)]
    b = BatchedDenseCRF2D(shapes, L)
    b.addPairwiseGaussian(sxy=3, compat=3)
    b.addPairwiseBilateral(sxy=30, srgb=13, rgbims=[img for img, _, _ in cases], compat=10)
    for i in (0, 1):
        b.model.set_unary_instance(i, cases[i][1])
    Qb = b.inference(n)

    def alone(i, U):
        (w, h), img = shapes[i], cases[i][0]
        s = capi.Model.create_2d(w, h, L)
        s.set_unary(U)
        s.add_gaussian(3.0, 3.0, 3.0)
        s.add_bilateral(30.0, 30.0, 13.0, 13.0, 13.0, img, 10.0)
        return np.ascontiguousarray(s.inference(n))

    for i in (0, 1):
        assert np.array_equal(Qb[i].view(np.int32), alone(i, cases[i][1]).view(np.int32)), "instance %d" % i
    Z = np.zeros((L, shapes[2][0] * shapes[2][1]), np.float32)
    assert np.array_equal(Qb[2].view(np.int32), alone(2, Z).view(np.int32))
    # the full unary at once, then the third instance's own: all three as given
    full = np.concatenate([U for _, U, _ in cases], axis=1)
    b.model.set_unary(full)
    b.model.set_unary_instance(2, cases[2][1])
    Qb = b.inference(n)
    for i in range(3):
        assert np.array_equal(Qb[i].view(np.int32), alone(i, cases[i][1]).view(np.int32)), "instance %d" % i


@pytest.mark.parametrize("L", [3, 130])
def test_batched_tiny_and_wide_instances(L):
    """A batch mixing a one-pixel image, a one-row image and ordinary ones, at a label
    count that fits a wave's lanes (3) and one that needs block-wide groups (130): each
    instance bit-identical to its image run alone, and within 1e-5 of the oracle."""
    shapes = [(1, 1), (29, 1), (12, 9), (40, 31)]
    n = 4
    cases = [image_case(w, h, L, seed=900 + i, sites=3) for i, (w, h) in enumerate(shapes)]
    b = BatchedDenseCRF2D(shapes, L)
    b.setUnaryEnergy([U for _, U, _ in cases])
    b.addPairwiseGaussian(sxy=3, compat=3)
    b.addPairwiseBilateral(sxy=30, srgb=13, rgbims=[img for img, _, _ in cases], compat=10)
    Qb = b.inference(n)
    for i, ((w, h), (img, U, _)) in enumerate(zip(shapes, cases)):
        s = capi.Model.create_2d(w, h, L)
        s.set_unary(U)
        s.add_gaussian(3.0, 3.0, 3.0)
        s.add_bilateral(30.0, 30.0, 13.0, 13.0, 13.0, img, 10.0)
        Qs = s.inference(n)
        assert np.array_equal(Qb[i].view(np.int32), np.ascontiguousarray(Qs).view(np.int32)), "instance %d" % i
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


def _volume_case(dims, L, seed):
    rng = np.random.default_rng(seed)
    feats_g, feats_b, Us, gts, sizes = [], [], [], [], []
    for (Z, Y, X) in dims:
        zz, yy, xx = np.meshgrid(np.arange(Z), np.arange(Y), np.arange(X), indexing="ij")
        pos = np.stack([zz, yy, xx]).reshape(3, -1).astype(np.float32)
        inten = rng.uniform(0, 1, (1, Z * Y * X)).astype(np.float32)
        feats_g.append(pos / np.float32(3))
        feats_b.append(np.concatenate([pos / np.float32(4), inten / np.float32(0.1)]))
        Us.append(rng.standard_normal((L, Z * Y * X)).astype(np.float32))
        gts.append(rng.integers(0, L, Z * Y * X))
        sizes.append(Z * Y * X)
    return feats_g, feats_b, Us, gts, sizes


def test_batched_volume_gradients_match_single():
    """Forward+backward through the autograd layer on a batch of volumes (the C4 shape:
    d = 3 Gaussian and d = 4 bilateral features, Diagonal compatibilities, L = 4) equals
    the per-volume passes: dU bit for bit (every per-voxel kernel of the batch is the
    per-volume kernel), compatibility and term-weight gradients (sums over voxels,
    grouped differently) to 1e-5 relative."""
    import torch
    from pydensecrf_amd.autograd import CRFLayer
    L, n = 4, 5
    dims = [(6, 9, 10), (8, 7, 6), (5, 10, 12)]
    feats_g, feats_b, Us, gts, sizes = _volume_case(dims, L, 31)
    w1 = np.array([1, 2, 2, 3], np.float32)
    w2 = np.array([2, 1, 3, 1], np.float32)

    def run(model, U, gt):
        layer = CRFLayer(model, n, [w1, w2], kweights=[1.5, 0.5])
        Ut = torch.tensor(np.ascontiguousarray(U.T), device="cuda", requires_grad=True)
        Q = layer(Ut)
        gtt = torch.tensor(gt, device="cuda")
        loss = -torch.log(Q.gather(1, gtt[:, None])).sum()
        loss.backward()
        return (Q.detach().cpu().numpy(), Ut.grad.cpu().numpy(), layer.compat.grad.cpu().numpy(),
                layer.kweights.grad.cpu().numpy())

    b = BatchedDenseCRF(sizes, L)
    b.setUnaryEnergy(Us)
    b.addPairwiseEnergy(feats_g, compat=w1)
    b.addPairwiseEnergy(feats_b, compat=w2)
    Qb, dUb, dcb, dkb = run(b.model, np.concatenate(Us, axis=1), np.concatenate(gts))
    off = np.concatenate([[0], np.cumsum(sizes)])
    dc_sum, dk_sum = 0.0, 0.0
    for i in range(len(dims)):
        m = capi.Model.create(sizes[i], L)
        m.add_pairwise(feats_g[i], w1)
        m.add_pairwise(feats_b[i], w2)
        Qs, dUs, dcs, dks = run(m, Us[i], gts[i])
        assert np.array_equal(Qb[off[i]:off[i + 1]].view(np.int32), Qs.view(np.int32)), i
        assert np.array_equal(dUb[off[i]:off[i + 1]].view(np.int32), dUs.view(np.int32)), i
        dc_sum = dc_sum + dcs.astype(np.float64)
        dk_sum = dk_sum + dks.astype(np.float64)
    np.testing.assert_allclose(dcb, dc_sum, rtol=1e-5, atol=1e-6 * np.abs(dc_sum).max())
    np.testing.assert_allclose(dkb, dk_sum, rtol=1e-5, atol=1e-6 * np.abs(dk_sum).max())


def test_batched_kernel_gradients_match_single():
    """Kernel-parameter gradients (learnable DIAG feature scales, set away from one so the
    lattices are rebuilt) of a batch of volumes equal the sum of the per-volume ones: the
    per-instance feature tables and lattices of the batch give each instance its own
    parameter-gradient sums."""
    import torch
    from pydensecrf_amd.autograd import CRFLayer
    L, n = 4, 4
    dims = [(6, 9, 10), (8, 7, 6), (5, 10, 12)]
    feats_g, feats_b, Us, gts, sizes = _volume_case(dims, L, 37)
    w1 = np.array([1, 2, 2, 3], np.float32)
    w2 = np.array([2, 1, 3, 1], np.float32)
    kp = np.array([1.1, 0.9, 1.05, 0.8, 1.2, 0.95, 1.0], np.float32)

    def run(model, U, gt):
        layer = CRFLayer(model, n, [w1, w2], kparams=kp, learn_kernel=True)
        Ut = torch.tensor(np.ascontiguousarray(U.T), device="cuda", requires_grad=True)
        Q = layer(Ut)
        gtt = torch.tensor(gt, device="cuda")
        loss = -torch.log(Q.gather(1, gtt[:, None])).sum()
        loss.backward()
        return Q.detach().cpu().numpy(), layer.kparams.grad.cpu().numpy().astype(np.float64)

    b = BatchedDenseCRF(sizes, L)
    b.setUnaryEnergy(Us)
    b.addPairwiseEnergy(feats_g, compat=w1)
    b.addPairwiseEnergy(feats_b, compat=w2)
    Qb, dkb = run(b.model, np.concatenate(Us, axis=1), np.concatenate(gts))
    off = np.concatenate([[0], np.cumsum(sizes)])
    dk_sum = 0.0
    for i in range(len(dims)):
        m = capi.Model.create(sizes[i], L)
        m.add_pairwise(feats_g[i], w1)
        m.add_pairwise(feats_b[i], w2)
        Qs, dks = run(m, Us[i], gts[i])
        np.testing.assert_array_equal(Qb[off[i]:off[i + 1]], Qs)
        dk_sum = dk_sum + dks
    np.testing.assert_allclose(dkb, dk_sum, rtol=1e-4, atol=1e-5 * np.abs(dk_sum).max())


def test_graph_replay_follows_unary_and_weights():
    """A captured graph replays with the model's current unary and term-weight values: a
    new unary (uploaded before the replay) and new weights in the same buffer give the
    bits of an eager inference; weights removed since the capture are an error."""
    import torch
    L, n = 21, 4
    cases = _cases(L)
    shapes = SHAPES[:3]
    b = BatchedDenseCRF2D(shapes, L)
    b.setUnaryEnergy([U for _, U, _ in cases[:3]])
    b.addPairwiseGaussian(sxy=3, compat=3)
    b.addPairwiseBilateral(sxy=30, srgb=13, rgbims=[img for img, _, _ in cases[:3]], compat=10)
    b.model.set_kernel_weights(np.array([1.0, 1.0], np.float32))
    Q = torch.zeros(b.N * L, device="cuda", dtype=torch.float32)
    b.capture(n, Q.data_ptr())
    rng = np.random.default_rng(5)
    newU = [(U + rng.normal(0, 0.5, U.shape)).astype(np.float32) for _, U, _ in cases[:3]]
    b.setUnaryEnergy(newU)
    b.model.set_kernel_weights(np.array([0.5, 2.0], np.float32))
    Q.fill_(-1.0)
    b.replay()
    torch.cuda.synchronize()
    got = Q.view(b.N, L).cpu().numpy().T
    ref = np.concatenate(b.inference(n), axis=1)
    assert np.array_equal(got.view(np.int32), ref.view(np.int32))
    b.model.set_kernel_weights(None)
    with pytest.raises(RuntimeError, match="capture again"):
        b.replay()
