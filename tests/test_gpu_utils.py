"""Device forms of the reference's helpers and debug energies against the reference:
utils.unary_from_softmax / unary_from_labels (pydensecrf/utils.py:6-79, here through the
numpy restatement that tests/test_api_cpu.py checks against the reference's own KAT)
and DenseCRF::unaryEnergy / pairwiseEnergy (densecrf.cpp:141-177) against the reference's
outputs in tests/golden/learning_small.npz."""
import os

import numpy as np
import pytest

from pydensecrf_amd import capi, utils

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("shape,scale,clip", [((21, 37, 29), None, 1e-5), ((5, 1000), 0.7, 1e-5),
                                               ((3, 64), None, None), ((33, 17, 3), 0.9, 1e-3)])
def test_unary_from_softmax_device(shape, scale, clip):
    import torch
    rng = np.random.default_rng(sum(shape))
    x = rng.standard_normal(shape).astype(np.float32) * 3
    sm = np.exp(x) / np.exp(x).sum(0, keepdims=True)
    sm[0].flat[::7] = 0.0  # clipped entries
    sm = sm.astype(np.float32)
    want = utils.unary_from_softmax(sm, scale, clip)  # (L, N)
    got = utils.unary_from_softmax_device(torch.tensor(sm, device="cuda"), scale, clip).cpu().numpy()
    if clip is None:
        ok = np.isfinite(want.T)
        np.testing.assert_allclose(got[ok], want.T[ok], rtol=2e-7, atol=2e-7)
        assert np.array_equal(np.isinf(got), np.isinf(want.T))
    else:
        np.testing.assert_allclose(got, want.T, rtol=2e-7, atol=2e-7)


@pytest.mark.parametrize("zero_unsure", [True, False])
def test_unary_from_labels_device(zero_unsure):
    import torch
    rng = np.random.default_rng(3)
    n = 7
    labels = rng.integers(0, n + 1 if zero_unsure else n, (23, 31))
    want = utils.unary_from_labels(labels, n, 0.7, zero_unsure)
    got = utils.unary_from_labels_device(torch.tensor(labels, device="cuda"), n, 0.7, zero_unsure).cpu().numpy()
    np.testing.assert_array_equal(got, want.T)


def test_debug_energies_reference_fixture():
    z = np.load(os.path.join(GOLDEN, "learning_small.npz"), allow_pickle=False)
    W, H, L = [int(v) for v in z["e_meta"]]
    g = capi.Model.create_2d(W, H, L)
    g.set_unary(z["e_U"])
    g.add_gaussian(3.0, 2.0, 3.0, ntype=3)
    g.add_bilateral(20.0, 30.0, 13.0, 17.0, 11.0, z["e_img"], np.arange(1, L + 1, dtype=np.float32), ntype=1)
    lab = z["e_labels"]
    np.testing.assert_array_equal(g.unary_energy(lab), z["e_unary"])
    for term, key in ((0, "e_pair0"), (1, "e_pair1"), (-1, "e_pair")):
        np.testing.assert_allclose(g.pairwise_energy(lab, term), z[key], rtol=1e-5, atol=1e-5)
