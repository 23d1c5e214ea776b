"""GPU adjoint mean-field: gradients against the reference and against FP64 finite
differences.

* dcrf_gradient (DenseCRF::gradient with LogLikelihood) against the reference's own
  outputs (tests/golden/crf_small.npz, written from the reference C++) and against the
  oracle on larger cases: objective within 1e-5 relative, unary and compatibility
  gradients within 1e-4 of their largest entry.
* CRFLayer (torch.autograd) against FP64 central differences (h = 1e-3) of the
  cross-entropy computed by the oracle on the same lattices: relative error <= 1e-3 on
  unary entries, every compatibility parameter and the term weights.
"""
import os

import numpy as np
import pytest

from oracle import pyoracle as po
from pydensecrf_amd import capi
from tests.helpers import image_case

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _rel_close(a, b, rel):
    scale = max(np.abs(b).max(), 1e-30)
    return np.abs(a - b).max() <= rel * scale


@pytest.mark.parametrize("i", range(8))
def test_gradient_reference_fixtures(i):
    z = np.load(os.path.join(GOLDEN, "crf_small.npz"), allow_pickle=False)
    W, H, L, ck, nt, n = [int(v) for v in z["c%d_meta" % i]]
    p = "c%d_" % i
    c1, c2 = z[p + "c1"], z[p + "c2"]
    if ck == 0:
        c1, c2 = float(c1[0]), float(c2[0])
    g = capi.Model.create_2d(W, H, L)
    g.set_unary(z[p + "U"])
    g.add_gaussian(3.0, 2.0, c1, ntype=nt)
    g.add_bilateral(20.0, 30.0, 13.0, 17.0, 11.0, z[p + "img"], c2, ntype=nt)
    obj, ug, cg, kg = g.gradient(n, z[p + "gt"])
    assert abs(obj - float(z[p + "obj"])) <= 1e-5 * max(1.0, abs(obj))
    assert _rel_close(ug, z[p + "unary_grad"], 1e-4)
    assert _rel_close(cg, z[p + "compat_grad"], 1e-4)
    np.testing.assert_array_equal(g.compat_params(), _packed(c1, c2, ck, L))


def _packed(c1, c2, ck, L):
    out = []
    for c in (c1, c2):
        if ck == 0:
            out.append([c])
        elif ck == 1:
            out.append(list(c))
        else:
            Wm = 0.5 * (np.asarray(c, np.float32) + np.asarray(c, np.float32).T)
            out.append([Wm[a, b] for a in range(L) for b in range(a, L)])
    return np.concatenate([np.asarray(o, np.float32) for o in out])


@pytest.mark.parametrize("ck", ["potts", "diag", "matrix"])
def test_gradient_vs_oracle_l21(ck):
    W, H, L, n = 72, 56, 21, 5
    img, U, gt = image_case(W, H, L, seed=33)
    rng = np.random.default_rng(1)
    if ck == "potts":
        c1, c2 = 3.0, 10.0
    elif ck == "diag":
        c1, c2 = rng.uniform(1, 3, L).astype(np.float32), rng.uniform(5, 10, L).astype(np.float32)
    else:
        c1 = (-3 * np.eye(L) + rng.normal(0, 0.1, (L, L))).astype(np.float32)
        c2 = (-10 * np.eye(L) + rng.normal(0, 0.1, (L, L))).astype(np.float32)
    g = capi.Model.create_2d(W, H, L)
    o = po.OracleCRF(W=W, H=H, L=L)
    for m in (g, o):
        m.set_unary(U)
        m.add_gaussian(3.0, 3.0, c1)
        m.add_bilateral(80.0, 80.0, 13.0, 13.0, 13.0, img, c2)
    obj, ug, cg, _ = g.gradient(n, gt)
    r, ugo, cgo = o.gradient(n, gt)
    assert abs(obj - r) <= 1e-5 * abs(r)
    assert _rel_close(ug, ugo, 1e-4)
    assert _rel_close(cg, cgo, 1e-4)


def test_kernel_weights_forward():
    """Term weights w_k scale each message: w = (2, 0.5) equals compat (2c1, 0.5c2)."""
    W, H, L = 40, 32, 5
    img, U, _ = image_case(W, H, L, seed=4)
    a = capi.Model.create_2d(W, H, L)
    b = capi.Model.create_2d(W, H, L)
    a.set_unary(U)
    b.set_unary(U)
    a.add_gaussian(3.0, 3.0, 3.0)
    a.add_bilateral(30.0, 30.0, 13.0, 13.0, 13.0, img, 10.0)
    b.add_gaussian(3.0, 3.0, 6.0)
    b.add_bilateral(30.0, 30.0, 13.0, 13.0, 13.0, img, 5.0)
    a.set_kernel_weights(np.array([2.0, 0.5], np.float32))
    np.testing.assert_allclose(a.inference(5), b.inference(5), atol=2e-6)


def _fd(fun, x, h):
    return (fun(x + h) - fun(x - h)) / (2 * h)


@pytest.mark.parametrize("ck", ["potts", "diag", "matrix"])
def test_autograd_central_differences(ck):
    import torch
    from pydensecrf_amd.autograd import CRFLayer

    W, H, L, n = 11, 9, 3, 4
    N = W * H
    img, U, gt = image_case(W, H, L, seed=12)
    rng = np.random.default_rng(2)
    U = (U * 0.5).astype(np.float32)
    if ck == "potts":
        c = [1.5, 2.5]
    elif ck == "diag":
        c = [rng.uniform(0.5, 2, L).astype(np.float32), rng.uniform(0.5, 2, L).astype(np.float32)]
    else:
        c = [(-np.eye(L) + rng.normal(0, 0.3, (L, L))).astype(np.float32),
             (-2 * np.eye(L) + rng.normal(0, 0.3, (L, L))).astype(np.float32)]
    kw0 = np.array([0.8, 1.3], np.float32)
    g = capi.Model.create_2d(W, H, L)
    g.add_gaussian(2.0, 2.0, c[0])
    g.add_bilateral(5.0, 5.0, 40.0, 40.0, 40.0, img, c[1])
    layer = CRFLayer(g, n, c, kweights=kw0)
    Ut = torch.tensor(np.ascontiguousarray(U.T), device="cuda", requires_grad=True)
    Q = layer(Ut)
    gtt = torch.tensor(gt.astype(np.int64), device="cuda")
    loss = -torch.log(Q[torch.arange(N, device="cuda"), gtt]).mean()
    loss.backward()
    dU = Ut.grad.cpu().numpy().T
    dc = layer.compat.grad.cpu().numpy()
    dk = layer.kweights.grad.cpu().numpy()

    o = po.OracleCRF(W=W, H=H, L=L)
    o.set_unary(U)
    o.add_gaussian(2.0, 2.0, c[0])
    o.add_bilateral(5.0, 5.0, 40.0, 40.0, 40.0, img, c[1])
    raw = np.concatenate([np.atleast_1d(np.asarray(x, np.float64)).ravel() for x in c])
    U64 = U.astype(np.float64)
    kw64 = kw0.astype(np.float64)
    ce = o.ce_fp64(n, gt, U64, raw, kw64)
    assert abs(float(loss) - ce) <= 1e-5 * abs(ce)
    h = 1e-3
    for (l, p) in [(0, 0), (1, 7), (2, 33), (0, 50), (1, 98)]:
        def f(v, l=l, p=p):
            Up = U64.copy()
            Up[l, p] = v
            return o.ce_fp64(n, gt, Up, raw, kw64)
        fd = _fd(f, U64[l, p], h)
        assert abs(dU[l, p] - fd) <= 1e-3 * max(abs(fd), 1e-4), (l, p, dU[l, p], fd)
    for i in range(raw.size):
        def f(v, i=i):
            r2 = raw.copy()
            r2[i] = v
            return o.ce_fp64(n, gt, U64, r2, kw64)
        fd = _fd(f, raw[i], h)
        assert abs(dc[i] - fd) <= 1e-3 * max(abs(fd), 1e-4), (i, dc[i], fd)
    for k in range(2):
        def f(v, k=k):
            k2 = kw64.copy()
            k2[k] = v
            return o.ce_fp64(n, gt, U64, raw, k2)
        fd = _fd(f, kw64[k], h)
        assert abs(dk[k] - fd) <= 1e-3 * max(abs(fd), 1e-4), (k, dk[k], fd)


def test_autograd_repeatable_and_guarded():
    import torch
    from pydensecrf_amd.autograd import CRFLayer

    W, H, L = 24, 20, 4
    img, U, gt = image_case(W, H, L, seed=3)
    g = capi.Model.create_2d(W, H, L)
    g.add_gaussian(3.0, 3.0, 3.0)
    g.add_bilateral(20.0, 20.0, 13.0, 13.0, 13.0, img, 10.0)
    layer = CRFLayer(g, 5, [3.0, 10.0])
    Ut = torch.tensor(np.ascontiguousarray(U.T), device="cuda", requires_grad=True)
    grads = []
    for _ in range(2):
        Ut.grad = None
        layer.zero_grad()
        Q = layer(Ut)
        (Q[:, 0].sum()).backward()
        grads.append((Ut.grad.clone(), layer.compat.grad.clone()))
    assert torch.equal(grads[0][0], grads[1][0]) and torch.equal(grads[0][1], grads[1][1])
    Q1 = layer(Ut)
    layer(Ut)  # a second forward invalidates the first graph's checkpoints
    with pytest.raises(RuntimeError):
        Q1.sum().backward()


@pytest.mark.parametrize("L", [4, 7, 130])
def test_gradient_label_shapes(L):
    W, H, n = 19, 13, 3
    img, U, gt = image_case(W, H, L, seed=L)
    rng = np.random.default_rng(L)
    c1 = (-np.eye(L) + rng.normal(0, 0.1, (L, L))).astype(np.float32)
    c2 = rng.uniform(0.5, 2, L).astype(np.float32)
    g = capi.Model.create_2d(W, H, L)
    o = po.OracleCRF(W=W, H=H, L=L)
    for m in (g, o):
        m.set_unary(U)
        m.add_gaussian(3.0, 3.0, c1)
        m.add_bilateral(10.0, 10.0, 13.0, 13.0, 13.0, img, c2)
    obj, ug, cg, _ = g.gradient(n, gt)
    r, ugo, cgo = o.gradient(n, gt)
    assert abs(obj - r) <= 1e-5 * abs(r)
    assert _rel_close(ug, ugo, 1e-4)
    assert _rel_close(cg, cgo, 1e-4)
