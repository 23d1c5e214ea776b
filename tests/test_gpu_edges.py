"""Degenerate shapes on the GPU against the oracle: one-pixel, one-row and one-column
images (pixel counts not divisible by 4, so the reference's SSE init pads lanes,
permutohedral.cpp:191-196), through inference, the gradient and the objectives; and the
softmax helper with label counts beyond one LDS tile row."""
import numpy as np
import pytest

from oracle import pyoracle as po
from pydensecrf_amd import capi, utils
from tests.helpers import image_case

pytestmark = pytest.mark.gpu
Q_TOL = 1e-5


def _pair(W, H, L, seed):
    img, U, gt = image_case(W, H, L, seed=seed, sites=2)
    g = capi.Model.create_2d(W, H, L)
    o = po.OracleCRF(W=W, H=H, L=L)
    for m in (g, o):
        m.set_unary(U)
        m.add_gaussian(3.0, 3.0, 3.0)
        m.add_bilateral(20.0, 20.0, 13.0, 13.0, 13.0, img, 10.0)
    return g, o, gt


@pytest.mark.parametrize("W,H", [(1, 1), (1, 17), (23, 1), (2, 3)])
@pytest.mark.parametrize("L", [3, 21])
def test_tiny_images(W, H, L):
    g, o, gt = _pair(W, H, L, seed=W * 100 + H)
    assert np.abs(g.inference(5) - o.inference(5)).max() <= Q_TOL
    obj, ug, cg, _ = g.gradient(3, gt)
    r, ugo, cgo = o.gradient(3, gt)
    assert abs(obj - r) <= 1e-5 * max(1.0, abs(r))
    assert np.abs(ug - ugo).max() <= 1e-4 * max(np.abs(ugo).max(), 1e-30) + 1e-7
    assert np.abs(cg - cgo).max() <= 1e-4 * max(np.abs(cgo).max(), 1e-30) + 1e-7
    for kind in (capi.OBJECTIVE_HAMMING, capi.OBJECTIVE_IOU):
        cw = np.ones(L, np.float32) if kind == capi.OBJECTIVE_HAMMING else None
        r1 = g.gradient_objective(2, gt, kind, 0.0, cw)[0]
        r2 = o.gradient_obj(2, gt, kind, 0.0, cw)[0]
        assert abs(r1 - r2) <= 1e-5 * max(1.0, abs(r2)), (kind, r1, r2)


@pytest.mark.parametrize("L", [33, 300])
def test_unary_from_softmax_many_labels(L):
    import torch
    rng = np.random.default_rng(L)
    x = rng.standard_normal((L, 9, 13)).astype(np.float32)
    sm = (np.exp(x) / np.exp(x).sum(0, keepdims=True)).astype(np.float32)
    want = utils.unary_from_softmax(sm, 0.8, 1e-5)
    got = utils.unary_from_softmax_device(torch.tensor(sm, device="cuda"), 0.8, 1e-5).cpu().numpy()
    np.testing.assert_allclose(got, want.T, rtol=2e-7, atol=2e-7)
