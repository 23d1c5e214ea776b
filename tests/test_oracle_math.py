"""The oracle against what the mathematics fixes, independently of the reference build.

Each check is chosen so that a plausible slip in the restatement (a dropped term, a
wrong sign or index, a transposed operand) fails it:
  lattice  barycentric weights sum to one and are non-negative; every vertex of
           remainder r has all d+1 coordinates = r mod (d+1) and coordinates summing to
           zero; sum_r b_r key_r reproduces the elevated point; blur neighbours are
           mutual; a brute-force search over nearby lattice points finds the same
           enclosing simplex for tiny d.
  filter   linearity; the reversed filter is the exact transpose of the forward one
           (FP64, dense matrix on tiny N); K*1 gives the normaliser.
  model    one step equals softmax(-U - sum_k compat_k(K~_k Q)) built from the oracle's
           per-term messages and the closed-form compatibilities; KL equals its closed
           form; the LogLikelihood gradient matches FP64 central differences.
  Eigen    the packet exp is within 2 ulp of exp; the packet sum is within n*eps of
           the exact sum.
"""
import itertools
import math

import numpy as np
import pytest

from oracle import pyoracle as po


def elevate(f):
    """E f of the permutohedral embedding (float64), (d+1, N)."""
    d, N = f.shape
    inv_std = math.sqrt(2.0 / 3.0) * (d + 1)
    scale = np.array([1.0 / math.sqrt((i + 2) * (i + 1)) * inv_std for i in range(d)])
    e = np.zeros((d + 1, N))
    sm = np.zeros(N)
    for j in range(d, 0, -1):
        cf = f[j - 1] * scale[j - 1]
        e[j] = sm - j * cf
        sm += cf
    e[0] = sm
    return e


def full_key(k):
    return np.concatenate([k, [-k.sum()]])


@pytest.mark.parametrize("d,N", [(1, 30), (2, 60), (3, 80), (5, 97), (8, 41)])
def test_lattice_invariants(d, N):
    rng = np.random.default_rng(10 + d)
    f = (rng.standard_normal((d, N)) * 2.5).astype(np.float32)
    lat = po.OracleLattice(f).arrays
    b = lat.bary[:N].astype(np.float64)
    assert np.all(b >= -1e-6)
    np.testing.assert_allclose(b.sum(1), 1.0, atol=2e-6)
    keys = lat.keys.astype(np.int64)
    el = elevate(f.astype(np.float64))
    for p in range(N):
        rec = np.zeros(d + 1)
        for r in range(d + 1):
            k = full_key(keys[lat.offset[p, r]])
            assert np.all(np.mod(k, d + 1) == r), "vertex of remainder %d has a coordinate off its class" % r
            rec += b[p, r] * k
        np.testing.assert_allclose(rec, el[:, p], atol=1e-4 * max(1.0, np.abs(el[:, p]).max()))
    # neighbour relation is mutual and matches the key arithmetic
    for j in range(d + 1):
        for v in range(lat.V):
            n1, n2 = lat.nbr[j, v]
            if n1 >= 0:
                assert lat.nbr[j, n1, 1] == v
                kv, k1 = full_key(keys[v]), full_key(keys[n1])
                step = -np.ones(d + 1, np.int64)
                step[j] = d
                assert np.array_equal(k1 - kv, step)
            if n2 >= 0:
                assert lat.nbr[j, n2, 0] == v


def test_enclosing_simplex_brute_force():
    """d = 2: among all points of the lattice (d+1)A*_d near the elevated point, the d+1
    vertices the oracle picked are exactly those of the simplex that contains it."""
    d = 2
    rng = np.random.default_rng(3)
    f = (rng.standard_normal((d, 40)) * 3).astype(np.float32)
    lat = po.OracleLattice(f).arrays
    el = elevate(f.astype(np.float64))
    for p in range(40):
        x = el[:, p]
        base = np.floor(x[:d] / (d + 1)).astype(int) * (d + 1)
        cands = []
        for off in itertools.product(range(-d - 1, 2 * (d + 1) + 1), repeat=d):
            k = base + np.array(off)
            kf = full_key(k)
            r = kf[0] % (d + 1)
            if np.all(np.mod(kf, d + 1) == r):
                cands.append(kf)
        cands = np.array(cands, dtype=np.float64)
        chosen = np.array([full_key(lat.keys[lat.offset[p, r]].astype(np.int64)) for r in range(d + 1)], float)
        # barycentric solve over the chosen simplex must be non-negative and exact
        A = np.vstack([chosen.T, np.ones(d + 1)])
        w = np.linalg.lstsq(A, np.concatenate([x, [1.0]]), rcond=None)[0]
        assert np.all(w >= -1e-9)
        # no candidate simplex from a different remainder-0 anchor is strictly closer:
        # the nearest remainder-0 point is a vertex of the chosen simplex
        zero_pts = cands[np.mod(cands[:, 0], d + 1) == 0]
        nearest = zero_pts[np.argmin(((zero_pts - x) ** 2).sum(1))]
        assert any(np.array_equal(nearest, c) for c in chosen)


def _dense(lat, N, reverse):
    K = np.zeros((N, N))
    for i in range(N):
        e = np.zeros((1, N))
        e[0, i] = 1.0
        K[:, i] = lat.compute(e, reverse=reverse, fp64=True)[0]
    return K


@pytest.mark.parametrize("d", [1, 2, 3, 5])
def test_filter_transpose_and_linearity(d):
    rng = np.random.default_rng(20 + d)
    N = 47
    f = (rng.standard_normal((d, N)) * 1.5).astype(np.float32)
    lat = po.OracleLattice(f)
    Kf = _dense(lat, N, False)
    Kr = _dense(lat, N, True)
    np.testing.assert_allclose(Kr, Kf.T, atol=1e-12)
    x, y = rng.standard_normal((3, N)), rng.standard_normal((3, N))
    lhs = lat.compute(2.0 * x - 0.5 * y, fp64=True)
    rhs = 2.0 * lat.compute(x, fp64=True) - 0.5 * lat.compute(y, fp64=True)
    np.testing.assert_allclose(lhs, rhs, atol=1e-12)
    # float path agrees with the FP64 dense matrix
    xf = x.astype(np.float32)
    np.testing.assert_allclose(lat.compute(xf), (Kf @ x.T).T, rtol=1e-5, atol=1e-5)


def test_norm_is_inverse_sqrt_of_K1():
    rng = np.random.default_rng(4)
    N, L = 120, 3
    f = (rng.standard_normal((3, N)) * 2).astype(np.float32)
    c = po.OracleCRF(N=N, L=L)
    c.set_unary(rng.standard_normal((L, N)).astype(np.float32))
    c.add_pairwise(f, 1.0)
    K1 = po.OracleLattice(f).compute(np.ones((1, N)), fp64=True)[0]
    np.testing.assert_allclose(c.norm(0), 1.0 / np.sqrt(K1), rtol=1e-5)


def _softmax(e):
    e = e - e.max(0, keepdims=True)
    x = np.exp(e)
    return x / x.sum(0, keepdims=True)


@pytest.mark.parametrize("ck", ["potts", "diag", "matrix"])
def test_step_closed_form(ck):
    rng = np.random.default_rng(5)
    W, H, L = 15, 11, 4
    N = W * H
    img = rng.integers(0, 256, (H, W, 3)).astype(np.uint8)
    U = rng.standard_normal((L, N)).astype(np.float32)
    if ck == "potts":
        c1, c2 = 3.0, 7.0
    elif ck == "diag":
        c1, c2 = rng.uniform(0.5, 2, L).astype(np.float32), rng.uniform(1, 5, L).astype(np.float32)
    else:
        c1 = rng.normal(0, 1, (L, L)).astype(np.float32)
        c2 = rng.normal(0, 1, (L, L)).astype(np.float32)
    o = po.OracleCRF(W=W, H=H, L=L)
    o.set_unary(U)
    o.add_gaussian(3.0, 3.0, c1)
    o.add_bilateral(8.0, 8.0, 20.0, 20.0, 20.0, img, c2)
    Q = o.start()
    np.testing.assert_allclose(Q, _softmax(-U.astype(np.float64)), atol=1e-6)
    # message of each term = compatibility applied to the symmetrically normalised filter
    xs = np.tile(np.arange(W), H).astype(np.float32)
    ys = np.repeat(np.arange(H), W).astype(np.float32)
    rgb = img.reshape(-1, 3).T.astype(np.float32)
    feats = [np.stack([xs / np.float32(3), ys / np.float32(3)]),
             np.concatenate([np.stack([xs / np.float32(8), ys / np.float32(8)]), rgb / np.float32(20)])]
    E = -U.astype(np.float64)
    for k, c in enumerate((c1, c2)):
        msg = o.apply(k, Q).astype(np.float64)
        nrm = o.norm(k).astype(np.float64)
        filt = po.OracleLattice(feats[k]).compute(Q * nrm, fp64=True) * nrm
        if ck == "potts":
            ref = -c * filt
        elif ck == "diag":
            ref = c[:, None] * filt
        else:
            ref = (0.5 * (c + c.T)).astype(np.float64) @ filt
        np.testing.assert_allclose(msg, ref, rtol=1e-4, atol=1e-5)
        E -= msg
    Q1, t1, t2 = o.step(Q)
    np.testing.assert_allclose(t1, E, rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(Q1, _softmax(E), atol=1e-6)


def test_kl_closed_form():
    rng = np.random.default_rng(6)
    W, H, L = 12, 10, 3
    N = W * H
    img = rng.integers(0, 256, (H, W, 3)).astype(np.uint8)
    U = rng.standard_normal((L, N)).astype(np.float32)
    o = po.OracleCRF(W=W, H=H, L=L)
    o.set_unary(U)
    o.add_gaussian(2.0, 2.0, 2.0)
    o.add_bilateral(6.0, 6.0, 30.0, 30.0, 30.0, img, 4.0)
    Q = o.inference(3).astype(np.float64)
    kl = (Q * np.log(np.maximum(Q, 1e-20))).sum() + (U * Q).sum()
    for k in range(2):
        kl += (Q * o.apply(k, Q.astype(np.float32))).sum()
    assert abs(o.kl(Q.astype(np.float32)) - kl) <= 1e-5 * abs(kl)


@pytest.mark.parametrize("ck", ["potts", "diag", "matrix"])
def test_gradient_central_differences(ck):
    """The float oracle's DenseCRF::gradient (of +mean log-likelihood) against FP64 central
    differences (h = 1e-3) of the cross-entropy on the same lattices."""
    rng = np.random.default_rng(8)
    W, H, L, n = 9, 7, 3, 3
    N = W * H
    img = rng.integers(0, 256, (H, W, 3)).astype(np.uint8)
    U = (rng.standard_normal((L, N)) * 0.5).astype(np.float32)
    gt = rng.integers(0, L, N).astype(np.int16)
    if ck == "potts":
        c1, c2 = 1.5, 2.5
        raw = np.array([1.5, 2.5])
    elif ck == "diag":
        c1, c2 = rng.uniform(0.5, 2, L).astype(np.float32), rng.uniform(0.5, 2, L).astype(np.float32)
        raw = np.concatenate([c1, c2]).astype(np.float64)
    else:
        c1 = rng.normal(0, 1, (L, L)).astype(np.float32)
        c2 = rng.normal(0, 1, (L, L)).astype(np.float32)
        raw = np.concatenate([c1.ravel(), c2.ravel()]).astype(np.float64)
    o = po.OracleCRF(W=W, H=H, L=L)
    o.set_unary(U)
    o.add_gaussian(2.0, 2.0, c1)
    o.add_bilateral(5.0, 5.0, 40.0, 40.0, 40.0, img, c2)
    r, ug, cg = o.gradient(n, gt)
    h = 1e-3
    ce0 = o.ce_fp64(n, gt)
    assert abs(-r - ce0) <= 1e-5 * abs(ce0)
    for (l, p) in [(0, 0), (1, 5), (2, 17), (0, 40), (2, 62)]:
        Up, Um = U.astype(np.float64), U.astype(np.float64)
        Up = Up.copy()
        Um = Um.copy()
        Up[l, p] += h
        Um[l, p] -= h
        fd = (o.ce_fp64(n, gt, U_LxN=Up) - o.ce_fp64(n, gt, U_LxN=Um)) / (2 * h)
        assert abs(-ug[l, p] - fd) <= 1e-3 * max(abs(fd), 1e-3), (l, p, ug[l, p], fd)
    # compatibility: FD over raw parameters, mapped to the reference's packing
    fd = np.zeros_like(raw)
    for i in range(raw.size):
        rp, rm = raw.copy(), raw.copy()
        rp[i] += h
        rm[i] -= h
        fd[i] = (o.ce_fp64(n, gt, compat_raw=rp) - o.ce_fp64(n, gt, compat_raw=rm)) / (2 * h)
    if ck == "matrix":
        # reference parameter (i, j>=i) drives both W(i,j) and W(j,i); W = (m + m^T)/2, so
        # dCE/dparam(i,j) = dCE/dm(i,j) + dCE/dm(j,i) for i != j, and dCE/dm(i,i) on the diagonal
        packed = []
        for t in range(2):
            G = fd[t * L * L:(t + 1) * L * L].reshape(L, L)
            for i in range(L):
                for j in range(i, L):
                    packed.append(G[i, j] + G[j, i] if i != j else G[i, i])
        fd = np.array(packed)
    np.testing.assert_allclose(-cg, fd, rtol=1e-3, atol=1e-5)


@pytest.mark.parametrize("d,N", [(2, 30), (3, 24), (5, 16)])
def test_lattice_gradient_central_differences(d, N):
    """Permutohedral::gradient(a, c) is the gradient of a^T K^T c (K^T: the reverse blur
    order) with respect to the features: central differences (h = 1e-3) of the FP64
    filter on lattices rebuilt at f +- h e_ij, generic feature positions."""
    rng = np.random.default_rng(40 + d)
    f = (rng.standard_normal((d, N)) * 2).astype(np.float32)
    a = rng.standard_normal((2, N))
    c = rng.standard_normal((2, N))
    df = po.OracleLattice(f).gradient(a.astype(np.float32), c.astype(np.float32))
    h = 1e-3
    fd = np.zeros((d, N))
    for j in range(d):
        for i in range(N):
            fp, fm = f.copy(), f.copy()
            fp[j, i] += h
            fm[j, i] -= h
            vp = (a * po.OracleLattice(fp).compute(c, True, fp64=True)).sum()
            vm = (a * po.OracleLattice(fm).compute(c, True, fp64=True)).sum()
            fd[j, i] = (vp - vm) / (2 * h)
    assert np.abs(df - fd).max() <= 2e-3 * np.abs(fd).max()


@pytest.mark.parametrize("kt", [1, 2])
@pytest.mark.parametrize("nt", [1, 2, 3])
@pytest.mark.parametrize("ck", ["potts", "matrix"])
def test_kernel_gradient_central_differences(kt, nt, ck):
    """DenseCRF::gradient's kernel_grad against FP64 central differences of the
    cross-entropy over the kernel parameters (setKernelParameters rebuilds the lattices).
    The reference differentiates with the transposed filter in place of the filter, so
    the check uses 1-D features, where both blur axes are the same operator and the
    lattice filter is symmetric; that pins the normaliser terms, the compatibility
    product and the sum over iterations."""
    rng = np.random.default_rng(8)
    N, L, n = 80, 3, 3
    U = (rng.standard_normal((L, N)) * 0.5).astype(np.float32)
    gt = rng.integers(0, L, N).astype(np.int16)
    f1 = (rng.standard_normal((1, N)) * 3).astype(np.float32)
    f2 = (rng.standard_normal((1, N)) * 2).astype(np.float32)
    o = po.OracleCRF(N=N, L=L)
    o.set_unary(U)
    o.add_pairwise(f1, 1.5 if ck == "potts" else rng.normal(0, 1, (L, L)).astype(np.float32), kt, nt)
    o.add_pairwise(f2, 2.5, kt, nt)
    o.set_kernel_params(np.array([1.1, 0.8], np.float32))
    p0 = o.kernel_params().astype(np.float64)
    _, _, _, kg = o.gradient(n, gt, kernel=True)
    h = 1e-3
    fd = np.zeros_like(p0)
    for i in range(p0.size):
        pp, pm = p0.copy(), p0.copy()
        pp[i] += h
        pm[i] -= h
        o.set_kernel_params(pp.astype(np.float32))
        cp = o.ce_fp64(n, gt)
        o.set_kernel_params(pm.astype(np.float32))
        cm = o.ce_fp64(n, gt)
        fd[i] = -(cp - cm) / (2 * h)
    np.testing.assert_allclose(kg, fd, rtol=2e-3, atol=2e-3 * np.abs(fd).max())


def test_kernel_gradient_no_normalization_central_differences():
    """NO_NORMALIZATION on 1-D features: the unnormalised messages make the objective
    steep, so the check differentiates the float objective itself over one iteration."""
    rng = np.random.default_rng(8)
    N, L = 80, 3
    U = (rng.standard_normal((L, N)) * 0.5).astype(np.float32)
    gt = rng.integers(0, L, N).astype(np.int16)
    f1 = (rng.standard_normal((1, N)) * 3).astype(np.float32)
    for n in (1, 2):
        o = po.OracleCRF(N=N, L=L)
        o.set_unary(U)
        o.add_pairwise(f1, 1.5, 1, po.NO_NORMALIZATION)
        _, _, _, kg = o.gradient(n, gt, kernel=True)
        h = 1e-3
        o.set_kernel_params(np.array([1 + h], np.float32))
        rp = o.gradient(n, gt)[0]
        o.set_kernel_params(np.array([1 - h], np.float32))
        rm = o.gradient(n, gt)[0]
        assert abs(kg[0] - (rp - rm) / (2 * h)) <= 2e-3 * abs(kg[0])


def test_exp_packet_accuracy():
    x = np.linspace(-87.0, 0.0, 20001).astype(np.float32)
    got = po.exp_packet(x).astype(np.float64)
    want = np.exp(x.astype(np.float64))
    ulp = np.spacing(want.astype(np.float32)).astype(np.float64)
    assert np.all(np.abs(got - want) <= 2 * ulp)


def test_sum_packet_accuracy():
    rng = np.random.default_rng(9)
    for n in [1, 3, 4, 7, 8, 21, 100, 1001]:
        x = rng.standard_normal(n).astype(np.float32)
        exact = math.fsum(x.astype(np.float64))
        assert abs(po.sum_packet(x) - exact) <= n * 6e-8 * np.abs(x).sum() + 1e-30
    # order: two accumulators over blocks of 8, a 4-block, the horizontal pair sum, the tail
    x = np.array([1e8, 1, -1e8, 1, 3, 0, 0, 0, 0, 0, 0, 0, 5], np.float32)
    p0 = [np.float32(1e8) + np.float32(0), np.float32(1) + np.float32(0), np.float32(-1e8) + np.float32(0),
          np.float32(1) + np.float32(0)]
    p0 = [p0[0] + np.float32(3), p0[1], p0[2], p0[3]]
    h0, h1 = np.float32(p0[0] + p0[2]), np.float32(p0[1] + p0[3])
    want = np.float32(np.float32(h0 + h1) + np.float32(5))
    assert po.sum_packet(x) == want
