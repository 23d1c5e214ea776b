"""Host-side behaviour of the drop-in API that needs no GPU.

Ports the reference suite's binding tests (pydensecrf/tests/test_eigen.py:7-41,
tests/test_utils.py:8-23, the error cases of pydensecrf/tests/test_dcrf.py:86-99 and
tests/issue29.py:25-36) to pydensecrf_amd, checks the C ABI's argument validation, and
that the library loads and exports every symbol include/dcrf_hip.h declares.
"""
import os
import re
import subprocess

import numpy as np
import pytest

import pydensecrf_amd.densecrf as dcrf
import pydensecrf_amd.eigen as e
from pydensecrf_amd import capi, utils

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


# ---------------------------------------------------------------- eigen buffers
def test_vector_conversion():
    v = np.random.randn(3).astype(np.float32)
    assert np.all(np.array(e.vectorXf(v)) == v)


def test_matrix_conversion():
    m = np.random.randn(3, 5).astype(np.float32)
    c = e.matrixXf(m)
    assert np.all(np.array(c) == m)
    # column-major storage, as the reference's Eigen copy
    assert np.array(c).strides == (4, 12)


def test_wrong_dims():
    with pytest.raises(ValueError):
        e.matrixXf(np.random.randn(3, 3, 3).astype(np.float32))


def test_wrong_type():
    with pytest.raises(ValueError):
        e.matrixXf(np.random.randn(3, 3).astype(np.float64))


def test_none_type():
    with pytest.raises(TypeError):
        e.matrixXf(None)


# ---------------------------------------------------------------- utils
def test_unary_kat():
    """tests/test_utils.py:10-23 of the reference."""
    M = 3
    U, P, N = 1. / M, 0.8, 0.2 / (M - 1)
    labels = np.array([[0, 1, 2, 3], [3, 2, 1, 0]])
    unary = -np.log(np.array([
        [U, P, N, N, N, N, P, U],
        [U, N, P, N, N, P, N, U],
        [U, N, N, P, P, N, N, U],
    ]))
    np.testing.assert_almost_equal(utils.compute_unary(labels, M, GT_PROB=P), unary)


def test_unary_from_softmax():
    sm = np.array([[0.2, 0.7], [0.8, 0.3]])
    np.testing.assert_allclose(utils.unary_from_softmax(sm), -np.log(sm).astype(np.float32))
    u = utils.unary_from_softmax(sm, scale=0.5)
    np.testing.assert_allclose(u, -np.log(0.5 * sm + 0.25).astype(np.float32), rtol=1e-6)


def test_pairwise_features_order_and_scale():
    f = utils.create_pairwise_gaussian(sdims=(2, 4), shape=(3, 5))
    assert f.shape == (2, 15) and f.dtype == np.float32
    # (y, x) order, 'ij' meshgrid: first feature is the row index / 2
    np.testing.assert_array_equal(f[0].reshape(3, 5), np.repeat(np.arange(3)[:, None], 5, 1) / np.float32(2))
    np.testing.assert_array_equal(f[1].reshape(3, 5), np.repeat(np.arange(5)[None], 3, 0) / np.float32(4))
    img = np.arange(3 * 5 * 3, dtype=np.uint8).reshape(3, 5, 3)
    b = utils.create_pairwise_bilateral(sdims=(2, 4), schan=(1, 2, 3), img=img, chdim=2)
    assert b.shape == (5, 15)
    np.testing.assert_array_equal(b[2:], (img.reshape(-1, 3).T / np.array([[1], [2], [3]], np.float32)))
    g = utils.create_pairwise_bilateral(sdims=(1, 1, 1), schan=0.5, img=np.ones((2, 3, 4)), chdim=-1)
    assert g.shape == (4, 24)
    np.testing.assert_array_equal(g[3], 2.0)


@pytest.mark.parametrize("sxy,srgb", [((3.0, 3.0), (13.0, 13.0, 13.0)), ((80.0, 80.0), (13.0, 13.0, 13.0)),
                                      ((2.5, 7.3), (13.0, 7.1, 0.3))])
def test_pairwise_features_match_densecrf2d(sxy, srgb):
    """The utils features are DenseCRF2D's own with the two coordinate rows swapped
    ((y, x) against (x, y), densecrf.cpp:61-81, float32 divisions): the oracle (bit-exact
    with the reference's lattice) builds the identical lattice from either."""
    from oracle import pyoracle as po
    W, H = 23, 17
    rng = np.random.default_rng(5)
    img = rng.integers(0, 256, (H, W, 3)).astype(np.uint8)
    sx, sy = sxy
    fb = utils.create_pairwise_bilateral(sdims=(sy, sx), schan=srgb, img=img, chdim=2)
    fg = utils.create_pairwise_gaussian(sdims=(sy, sx), shape=(H, W))
    assert fb.shape == (5, W * H) and fg.shape == (2, W * H)
    assert np.array_equal(fb[:2], fg)
    swap = [1, 0, 2, 3, 4]
    a = po.OracleCRF(W=W, H=H, L=2)
    a.add_gaussian(sx, sy, 1.0)
    a.add_bilateral(sx, sy, srgb[0], srgb[1], srgb[2], img, 1.0)
    b = po.OracleCRF(N=W * H, L=2)
    b.add_pairwise(np.ascontiguousarray(fg[[1, 0]]), 1.0)
    b.add_pairwise(np.ascontiguousarray(fb[swap]), 1.0)
    for k in range(2):
        la, lb = a.lattice(k), b.lattice(k)
        assert la.V == lb.V
        assert np.array_equal(la.offset, lb.offset)
        assert np.array_equal(la.bary.view(np.int32), lb.bary.view(np.int32))


def test_pairwise_features_nd_and_channel_axes():
    """A 3-D grid with the channel axis first, last, or absent; scales given for fewer
    channels than there are leave the rest unscaled."""
    vol = np.random.default_rng(1).integers(0, 9, (4, 5, 6)).astype(np.float32)
    f = utils.create_pairwise_bilateral(sdims=(1, 2, 3), schan=0.5, img=vol, chdim=-1)
    assert f.shape == (4, 120) and f.dtype == np.float32
    zz, yy, xx = np.meshgrid(np.arange(4), np.arange(5), np.arange(6), indexing="ij")
    np.testing.assert_array_equal(f[0], zz.ravel())
    np.testing.assert_array_equal(f[2], xx.ravel().astype(np.float32) / np.float32(3))
    np.testing.assert_array_equal(f[3], vol.ravel() / np.float32(0.5))
    two = np.stack([vol, vol + 1])  # channel axis 0
    g = utils.create_pairwise_bilateral(sdims=(1, 1, 1), schan=(2.0,), img=two, chdim=0)
    assert g.shape == (5, 120)
    np.testing.assert_array_equal(g[3], vol.ravel() / np.float32(2))
    np.testing.assert_array_equal(g[4], vol.ravel() + 1)
    h = utils.create_pairwise_bilateral(sdims=(1, 1, 1), schan=(2.0, 4.0), img=np.moveaxis(two, 0, -1), chdim=3)
    np.testing.assert_array_equal(h, np.concatenate([g[:4], (vol.ravel()[None] + 1) / np.float32(4)]))
    with pytest.raises(IndexError):
        utils.create_pairwise_gaussian(sdims=(1, 1, 1), shape=(3, 4))


def test_unary_from_labels_cases():
    """zero_unsure on and off, label maps of any shape, float32 energies of the double
    -log values, and the errors of out-of-range or non-integer labels."""
    lab = np.array([[0, 2], [3, 1]])
    U = utils.unary_from_labels(lab, 3, 0.7)
    assert U.dtype == np.float32 and U.shape == (3, 4)
    e_pos, e_neg, e_uni = np.float32(-np.log(0.7)), np.float32(-np.log(0.15)), np.float32(-np.log(1 / 3))
    np.testing.assert_array_equal(U[:, 0], [e_uni] * 3)
    np.testing.assert_array_equal(U[:, 1], [e_neg, e_pos, e_neg])
    np.testing.assert_array_equal(U[:, 2], [e_neg, e_neg, e_pos])
    V = utils.unary_from_labels(np.array([0, 2]), 3, 0.7, zero_unsure=False)
    np.testing.assert_array_equal(V[:, 0], [e_pos, e_neg, e_neg])
    with pytest.raises(IndexError):
        utils.unary_from_labels(np.array([4]), 3, 0.7)
    with pytest.raises(IndexError):
        utils.unary_from_labels(np.array([1.0]), 3, 0.7)
    with pytest.raises(AssertionError):
        utils.unary_from_labels(lab, 3, 1.0)


def test_unary_from_softmax_precision():
    """float32 probabilities: the log in float32 (clipped at float32(clip)); mixed with the
    uniform share: in double, rounded once."""
    sm = np.random.default_rng(2).dirichlet(np.ones(4), size=(3, 5)).transpose(2, 0, 1).astype(np.float32)
    sm[0, 0, 0] = 0.0
    u = utils.unary_from_softmax(sm)
    assert u.shape == (4, 15) and u.dtype == np.float32
    np.testing.assert_array_equal(u, -np.log(np.maximum(sm, np.float32(1e-5))).reshape(4, -1))
    v = utils.unary_from_softmax(sm, scale=0.8, clip=None)
    exp = -np.log((0.8 * sm).astype(np.float64) + 0.2 / 4)
    np.testing.assert_array_equal(v, exp.reshape(4, -1).astype(np.float32))
    with pytest.raises(AssertionError):
        utils.unary_from_softmax(sm, scale=1.5)


# ---------------------------------------------------------------- API errors
def test_compat_wrong_dtype():
    """pydensecrf/tests/test_dcrf.py:86-99: a float64 compatibility raises ValueError."""
    d = dcrf.DenseCRF2D(10, 10, 2)
    d.setUnaryEnergy(np.ones((2, 10 * 10), dtype=np.float32))
    with pytest.raises(ValueError):
        d.addPairwiseBilateral(sxy=(3, 3), srgb=(3, 3, 3), rgbim=np.zeros((10, 10, 3), np.uint8),
                               compat=np.array([1.0, 2.0]))


def test_compat_three_dims():
    d = dcrf.DenseCRF2D(4, 4, 2)
    with pytest.raises(ValueError):
        d.addPairwiseGaussian(3, np.zeros((2, 2, 2), np.float32))


def test_wrong_unary_shape():
    """tests/issue29.py:25-36: a unary of the wrong shape raises instead of crashing."""
    d = dcrf.DenseCRF(69 * 51 * 72, 5)
    with pytest.raises(ValueError):
        d.setUnaryEnergy(np.ones((5, 68 * 50), np.float32))


def test_wrong_feature_shape():
    d = dcrf.DenseCRF(100, 2)
    with pytest.raises(ValueError):
        d.addPairwiseEnergy(np.ones((2, 99), np.float32), compat=3)


def test_wrong_image_shape():
    d = dcrf.DenseCRF2D(10, 8, 2)
    with pytest.raises(ValueError):
        d.addPairwiseBilateral(sxy=3, srgb=3, rgbim=np.zeros((10, 8, 3), np.uint8), compat=3)
    with pytest.raises(ValueError):
        d.addPairwiseBilateral(sxy=3, srgb=3, rgbim=np.zeros((8, 10, 4), np.uint8), compat=3)


def test_none_arguments():
    d = dcrf.DenseCRF(10, 2)
    with pytest.raises(TypeError):
        d.setUnaryEnergy(None)
    with pytest.raises(TypeError):
        d.addPairwiseEnergy(None, compat=3)
    # Q buffers: a TypeError before any device work (not a read through None)
    from pydensecrf_amd import eigen
    q = eigen.MatrixXf()
    for call in (lambda: d.stepInference(None, q, q), lambda: d.stepInference(q, None, q),
                 lambda: d.klDivergence(None), lambda: d.currentMap(None)):
        with pytest.raises(TypeError):
            call()


def test_enums_match_reference_values():
    assert (dcrf.NO_NORMALIZATION, dcrf.NORMALIZE_BEFORE, dcrf.NORMALIZE_AFTER, dcrf.NORMALIZE_SYMMETRIC) == (0, 1, 2, 3)
    assert (dcrf.CONST_KERNEL, dcrf.DIAG_KERNEL, dcrf.FULL_KERNEL) == (0, 1, 2)


def test_step_api_shapes_without_device():
    d = dcrf.DenseCRF(12, 3)
    # no unary: startInference returns zeros without touching a device (densecrf.cpp:180-184)
    Q, t1, t2 = d.startInference()
    assert np.array(Q).shape == (3, 12) and np.all(np.array(Q) == 0)
    assert np.array(t1).shape == (0, 0)


def test_abi_validation():
    m = capi.Model.create(10, 3)
    with pytest.raises(ValueError):
        m.add_pairwise(np.zeros((16, 10), np.float32), 1.0)  # d > 15
    with pytest.raises(ValueError):
        capi.Model.create(10, capi.MAX_LABELS + 1)
    with pytest.raises(ValueError):
        capi.Model.create(0, 3)
    h = capi._P()
    assert capi.lib().dcrf_create_2d(0, 5, 2, capi.ctypes.byref(h)) == capi.ERR_VALUE
    assert b"positive" in capi.lib().dcrf_last_error()
    g = capi.Model.create(10, 3)
    assert capi.lib().dcrf_add_gaussian_2d(g.h, 1.0, 1.0, 0, capi.fptr(np.ones(1, np.float32)), 1, 3) == capi.ERR_STATE


@pytest.mark.skipif(capi.lib().dcrf_device_count() > 0, reason="checks the no-device error path")
def test_no_device_error_is_runtime_error():
    d = dcrf.DenseCRF2D(4, 4, 2)
    d.setUnaryEnergy(np.ones((2, 16), np.float32))
    with pytest.raises(RuntimeError):
        d.addPairwiseGaussian(3, 3)
    with pytest.raises(RuntimeError):
        d.inference(1)


# ---------------------------------------------------------------- the C ABI itself
def _declared_functions():
    src = open(os.path.join(ROOT, "include", "dcrf_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(dcrf_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    names = _declared_functions()
    assert len(names) >= 25
    lib = capi.lib()
    for n in names:
        assert hasattr(lib, n), n
    out = subprocess.run(["nm", "-D", "--defined-only", capi.LIB_PATH], stdout=subprocess.PIPE, text=True).stdout
    exported = set(re.findall(r" T (dcrf_[a-z0-9_]+)", out))
    assert set(names) <= exported
    # the ctypes binding types every declared function
    assert set(names) == set(capi.SIGNATURES)


def test_cython_module_links_the_library():
    import pydensecrf_amd.densecrf as mod
    out = subprocess.run(["ldd", mod.__file__], stdout=subprocess.PIPE, text=True).stdout
    assert "libdcrf_hip.so" in out
