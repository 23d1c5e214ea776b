"""Shared builders for the parity tests: the same model on the GPU (through the C ABI)
and on the oracle (oracle/liboracle.so, a restatement of the reference C++)."""
import numpy as np

from oracle import pyoracle as po
from pydensecrf_amd import capi


def image_case(W, H, L, seed=0, sites=5):
    from pydensecrf_amd import synthetic as syn
    img, U, gt = syn.make_image(W, H, L=L, sites=sites, seed=seed)
    return img, U, gt


def build_both(W, H, L, U, terms, generic=False, tile=None):
    """terms: list of ('gauss', sx, sy, compat, ntype) / ('bilat', sx, sy, sr, sg, sb, img, compat, ntype) /
    ('feat', features (d,N), compat, ntype).  tile: splat tile side of the GPU model."""
    g = capi.Model.create(W * H, L) if generic else capi.Model.create_2d(W, H, L)
    if tile is not None:
        g.set_splat_tile(tile)
    o = po.OracleCRF(N=W * H, L=L) if generic else po.OracleCRF(W=W, H=H, L=L)
    if U is not None:
        g.set_unary(U)
        o.set_unary(U)
    for t in terms:
        kind = t[0]
        if kind == "gauss":
            _, sx, sy, c, nt = t
            g.add_gaussian(sx, sy, c, ntype=nt)
            o.add_gaussian(sx, sy, c, ntype=nt)
        elif kind == "bilat":
            _, sx, sy, sr, sg, sb, img, c, nt = t
            g.add_bilateral(sx, sy, sr, sg, sb, img, c, ntype=nt)
            o.add_bilateral(sx, sy, sr, sg, sb, img, c, ntype=nt)
        else:
            _, f, c, nt = t
            g.add_pairwise(f, c, ntype=nt)
            o.add_pairwise(f, c, ntype=nt)
    return g, o


def assert_lattice_equal(glat, olat, N):
    d = olat.d
    assert glat["V"] == olat.V, (glat["V"], olat.V)
    assert np.array_equal(glat["offset"], olat.offset[:N]), "vertex offsets differ"
    assert np.array_equal(glat["rank"], olat.rank[:N]), "ranks differ"
    assert np.array_equal(glat["bary"].view(np.int32), olat.bary[:N].view(np.int32)), "barycentric bits differ"
    assert np.array_equal(glat["nbr"], olat.nbr), "blur neighbours differ"
    assert np.array_equal(glat["keys"], olat.keys), "vertex keys differ"
    assert glat["d"] == d


# FD bars per parameter group: |g - fd| <= 1e-3 max(|fd|, floor * max |fd| of the group).
# A unary entry is a difference of per-pixel terms of the size of the largest entries
# (sumAndNormalize), so FP32 resolves it to ~1e-6 of the largest: entries below 1% of
# it are held to 1e-5 of the largest (the reference's own gradient meets this at 5e-4).
# Compatibility entries are sums over all pixels (floor 1e-4); the term weights are two
# sums of the same kind (no floor).
FD_FLOOR = {"unary": 1e-2, "compat": 1e-4, "kweights": 0.0}


def fd_relative_errors(g_unary, g_pairs, g_kw, z):
    out = []
    for g, fd, fl in ((g_unary, z["fd_unary"], FD_FLOOR["unary"]), (g_pairs, z["fd_pairs"], FD_FLOOR["compat"]),
                      (g_kw, z["fd_kweights"], FD_FLOOR["kweights"])):
        out.append(np.abs(g - fd) / np.maximum(np.abs(fd), fl * np.abs(fd).max()))
    return out


def rosenbrock(x):
    """Extended Rosenbrock function of float32 x (even length), value and gradient in
    double (the gradient returned as float32, as a float energy hands it back)."""
    x = x.astype(np.float64)
    g = np.zeros_like(x)
    f = 0.0
    for i in range(0, x.size, 2):
        t1 = 1.0 - x[i]
        t2 = 10.0 * (x[i + 1] - x[i] * x[i])
        g[i + 1] = 20.0 * t2
        g[i] = -2.0 * (x[i] * g[i + 1] + t1)
        f += t1 * t1 + t2 * t2
    return f, g.astype(np.float32)


def crf_learning_problem():
    """A small learning problem: 16x12 image, 3 labels, Diagonal compatibilities and DIAG
    kernel parameters of a Gaussian and a bilateral term (13 parameters)."""
    from pydensecrf_amd import synthetic as syn
    img, U, gt = syn.make_image(16, 12, L=3, sites=4, seed=61)
    return img, U, gt


def crf_learning_energy(crf, gt, n_iter=3, l2=1e-2, kernel=True):
    """E(x) = -LogLikelihood(Q_n) + l2 / 2 ||x||^2 over x = (compatibility parameters[,
    kernel parameters]) of an oracle or reference CRF (set, then DenseCRF::gradient)."""
    nc = crf.compat_params().size

    def fun(x):
        crf.set_compat_params(x[:nc])
        if kernel:
            crf.set_kernel_params(x[nc:])
        r, _, cg, kg = crf.gradient_obj(n_iter, gt, 0, 0.0, None, kernel=kernel)
        x64 = x.astype(np.float64)
        g = -np.concatenate([cg, kg] if kernel else [cg]).astype(np.float64) + l2 * x64
        return -r + 0.5 * l2 * float(x64 @ x64), g.astype(np.float32)
    return fun


def crf_learning_model(cls, **kw):
    """The learning problem's CRF built with cls (OracleCRF / RefCRF)."""
    img, U, gt = crf_learning_problem()
    crf = cls(W=16, H=12, L=3, **kw)
    crf.set_unary(U)
    crf.add_gaussian(3.0, 2.0, np.array([1.0, 2.0, 1.5], np.float32))
    crf.add_bilateral(20.0, 30.0, 13.0, 17.0, 11.0, img, np.array([3.0, 4.0, 5.0], np.float32))
    return crf, gt
