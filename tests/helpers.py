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
