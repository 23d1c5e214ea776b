"""The drop-in module (pydensecrf_amd.densecrf, the reference's Python API) on the GPU.

Covers the user-facing contract of pydensecrf/densecrf.pyx:49-140: DenseCRF2D with
addPairwiseGaussian / addPairwiseBilateral, generic DenseCRF with the utils features
through addPairwiseEnergy, inference, startInference then stepInference updating the
same Q buffer in place (densecrf.cpp:187-201), klDivergence, and np.argmax(Q, 0).
Bars: the reference suite's own 10x10 problem (tests/golden/simple_10x10.npz, written by
the reference C++) within 1e-5; Q, the step temporaries and KL against the oracle within
1e-5 (KL relative); MAP labels identical.
"""
import os

import numpy as np
import pytest

from oracle import pyoracle as po
from pydensecrf_amd import synthetic as syn
from pydensecrf_amd import utils

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _dcrf():
    import pydensecrf_amd.densecrf as dcrf
    return dcrf


def test_reference_suite_10x10():
    """test_dcrf.py:62-83 with the reference's own outputs: DenseCRF2D with a bilateral
    term, and DenseCRF with utils.create_pairwise_bilateral features."""
    dcrf = _dcrf()
    z = np.load(os.path.join(GOLDEN, "simple_10x10.npz"), allow_pickle=False)
    U, img = z["U"], z["img"]
    d = dcrf.DenseCRF2D(10, 10, 2)
    d.setUnaryEnergy(U)
    d.addPairwiseBilateral(sxy=2, srgb=2, rgbim=img, compat=3)
    Q2d = np.array(d.inference(10))
    assert Q2d.shape == (2, 100)
    assert np.abs(Q2d - z["Q2d"]).max() <= 1e-5
    g = dcrf.DenseCRF(100, 2)
    g.setUnaryEnergy(U)
    feats = utils.create_pairwise_bilateral(sdims=(2, 2), schan=2, img=img, chdim=2)
    assert np.array_equal(feats, z["feats"])
    g.addPairwiseEnergy(feats, compat=3)
    Qg = np.array(g.inference(10))
    assert np.abs(Qg - z["Qgeneric"]).max() <= 1e-5
    assert np.array_equal(Q2d.argmax(0), Qg.argmax(0))  # the suite's own assertion


@pytest.mark.parametrize("generic", [False, True])
def test_step_inference_in_place(generic):
    """startInference, then stepInference mutating the same Q (seen through one
    np.asarray view) step after step, tmp1 / tmp2 as the reference leaves them,
    klDivergence of each Q, and the MAP of the last."""
    dcrf = _dcrf()
    W, H, L = 64, 48, 5
    img, U, _ = syn.make_image(W, H, L=L, sites=6, seed=77)
    o = po.OracleCRF(N=W * H, L=L) if generic else po.OracleCRF(W=W, H=H, L=L)
    o.set_unary(U)
    if generic:
        d = dcrf.DenseCRF(W * H, L)
        fg = utils.create_pairwise_gaussian(sdims=(3, 3), shape=(H, W))
        fb = utils.create_pairwise_bilateral(sdims=(50, 50), schan=(13, 13, 13), img=img, chdim=2)
        d.setUnaryEnergy(U)
        d.addPairwiseEnergy(fg, compat=3)
        d.addPairwiseEnergy(fb, compat=np.arange(1, L + 1, dtype=np.float32), kernel=dcrf.DIAG_KERNEL,
                            normalization=dcrf.NORMALIZE_SYMMETRIC)
        o.add_pairwise(fg, 3.0)
        o.add_pairwise(fb, np.arange(1, L + 1, dtype=np.float32))
    else:
        d = dcrf.DenseCRF2D(W, H, L)
        d.setUnaryEnergy(U)
        d.addPairwiseGaussian(sxy=3, compat=3)
        d.addPairwiseBilateral(sxy=50, srgb=13, rgbim=img, compat=np.arange(1, L + 1, dtype=np.float32))
        o.add_gaussian(3.0, 3.0, 3.0)
        o.add_bilateral(50.0, 50.0, 13.0, 13.0, 13.0, img, np.arange(1, L + 1, dtype=np.float32))
    Q, tmp1, tmp2 = d.startInference()
    view = np.asarray(Q)
    assert view.shape == (L, W * H)
    Qo = o.start()
    assert np.abs(view - Qo).max() <= 1e-6
    for step in range(4):
        d.stepInference(Q, tmp1, tmp2)
        Qo, t1o, t2o = o.step(Qo)
        # the same buffer, updated in place
        assert np.abs(view - Qo).max() <= 1e-5, step
        assert np.abs(np.asarray(tmp1) - t1o).max() <= 1e-4 * max(1.0, np.abs(t1o).max()), step
        assert np.abs(np.asarray(tmp2) - t2o).max() <= 1e-4 * max(1.0, np.abs(t2o).max()), step
        kl, klo = d.klDivergence(Q), o.kl(Qo)
        assert abs(kl - klo) <= 1e-5 * abs(klo), (step, kl, klo)
    assert np.array_equal(np.argmax(Q, axis=0), np.argmax(Qo, axis=0))
    Qi = np.array(d.inference(4))
    assert np.abs(Qi - view).max() <= 1e-6
    # the reference C++ interface's map / currentMap (not in its Python binding)
    assert np.array_equal(d.currentMap(Q), np.argmax(Q, axis=0).astype(np.int16))
    assert np.array_equal(d.map(4), Qi.argmax(0).astype(np.int16))
    from pydensecrf_amd import eigen
    with pytest.raises(ValueError):
        d.currentMap(eigen.MatrixXf())  # shape (0, 0)
    with pytest.raises(TypeError):
        d.currentMap(None)


def test_debug_energies_reference_fixture():
    """DenseCRF::unaryEnergy / pairwiseEnergy (densecrf.cpp:141-177; the reference's Python
    binding wraps neither) through the drop-in classes, against the reference's outputs in
    tests/golden/learning_small.npz: unary bit-exact, pairwise within 1e-5."""
    dcrf = _dcrf()
    z = np.load(os.path.join(GOLDEN, "learning_small.npz"), allow_pickle=False)
    W, H, L = [int(v) for v in z["e_meta"]]
    d = dcrf.DenseCRF2D(W, H, L)
    d.setUnaryEnergy(z["e_U"])
    d.addPairwiseGaussian(sxy=(3, 2), compat=3, normalization=dcrf.NORMALIZE_SYMMETRIC)
    d.addPairwiseBilateral(sxy=(20, 30), srgb=(13, 17, 11), rgbim=np.ascontiguousarray(z["e_img"]),
                           compat=np.arange(1, L + 1, dtype=np.float32), normalization=dcrf.NORMALIZE_BEFORE)
    lab = z["e_labels"]
    np.testing.assert_array_equal(d.unaryEnergy(lab), z["e_unary"])
    for term, key in ((0, "e_pair0"), (1, "e_pair1"), (-1, "e_pair")):
        np.testing.assert_allclose(d.pairwiseEnergy(lab, term), z[key], rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(d.pairwiseEnergy(lab), z["e_pair"], rtol=1e-5, atol=1e-5)
    with pytest.raises(ValueError):
        d.unaryEnergy(lab[:-1])
