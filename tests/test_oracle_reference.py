"""The oracle (oracle/dcrf_oracle.c) against the reference's own C++ outputs.

Golden fixtures in tests/golden/ were written from the reference build by
tests/golden/make_golden.py.  The bar is bit-exact equality (lattice arrays, Q, step
temporaries, KL, LogLikelihood gradients) except where the reference's result depends
on Eigen's blocked GEMM order (MatrixCompatibility), which the oracle restates as a
plain sequential sum: there the bar is 2e-6 relative.  When oracle/_ref is built (it is
in the build container), the live reference is compared on further random cases.
"""
import hashlib
import os

import numpy as np
import pytest

from oracle import pyoracle as po

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.int32)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def test_lattice_fixtures_bitexact():
    z = _load("lattice_small.npz")
    for i in range(int(z["ncases"])):
        f = z["c%d_feat" % i]
        lat = po.OracleLattice(f).arrays
        assert lat.V == int(z["c%d_V" % i])
        assert np.array_equal(lat.offset, z["c%d_offset" % i])
        assert np.array_equal(lat.rank, z["c%d_rank" % i])
        assert np.array_equal(bits(lat.bary), bits(z["c%d_bary" % i]))
        assert np.array_equal(lat.nbr, z["c%d_nbr" % i])


def _crf_case(z, i):
    W, H, L, ck, nt, n = [int(v) for v in z["c%d_meta" % i]]
    p = "c%d_" % i
    c1, c2 = z[p + "c1"], z[p + "c2"]
    if ck == 0:
        c1, c2 = float(c1[0]), float(c2[0])
    o = po.OracleCRF(W=W, H=H, L=L)
    o.set_unary(z[p + "U"])
    o.add_gaussian(3.0, 2.0, c1, ntype=nt)
    o.add_bilateral(20.0, 30.0, 13.0, 17.0, 11.0, z[p + "img"], c2, ntype=nt)
    return o, (W, H, L, ck, nt, n), p


@pytest.mark.parametrize("i", range(8))
def test_crf_fixtures(i):
    z = _load("crf_small.npz")
    o, (W, H, L, ck, nt, n), p = _crf_case(z, i)
    Q = o.inference(n)
    Qs = o.start()
    Q1, t1, t2 = o.step(Qs)
    kl = o.kl(Q)
    r, ug, cg = o.gradient(n, z[p + "gt"])
    if ck != 2:
        assert np.array_equal(bits(Q), bits(z[p + "Q"]))
        assert np.array_equal(bits(Q1), bits(z[p + "Q1"]))
        assert np.array_equal(bits(t1), bits(z[p + "tmp1"]))
        assert np.array_equal(bits(t2), bits(z[p + "tmp2"]))
        assert kl == float(z[p + "kl"])
        assert r == float(z[p + "obj"])
        assert np.array_equal(bits(ug), bits(z[p + "unary_grad"]))
        assert np.array_equal(bits(cg), bits(z[p + "compat_grad"]))
    else:
        assert np.abs(Q - z[p + "Q"]).max() <= 2e-6
        assert np.abs(Q1 - z[p + "Q1"]).max() <= 2e-6
        np.testing.assert_allclose(t1, z[p + "tmp1"], rtol=2e-6, atol=2e-6)
        np.testing.assert_allclose(t2, z[p + "tmp2"], rtol=2e-6, atol=2e-6)
        assert abs(kl - float(z[p + "kl"])) <= 2e-6 * max(1.0, abs(kl))
        assert abs(r - float(z[p + "obj"])) <= 2e-6 * max(1.0, abs(r))
        scale = np.abs(z[p + "unary_grad"]).max()
        assert np.abs(ug - z[p + "unary_grad"]).max() <= 2e-5 * scale
        cs = np.abs(z[p + "compat_grad"]).max()
        assert np.abs(cg - z[p + "compat_grad"]).max() <= 2e-5 * cs


def test_lattice_gradient_fixtures_bitexact():
    """Permutohedral::gradient (permutohedral.cpp:612-695) on the reference's outputs."""
    z = _load("kgrad_small.npz")
    for i in range(int(z["nlat"])):
        df = po.OracleLattice(z["l%d_feat" % i]).gradient(z["l%d_a" % i], z["l%d_b" % i])
        assert np.array_equal(bits(df), bits(z["l%d_df" % i])), i


@pytest.mark.parametrize("i", range(8))
def test_kernel_gradient_fixtures(i):
    """DenseCRF::gradient's kernel_grad (pairwise.cpp:152-163, 202-207) after
    setKernelParameters: bit-exact for DIAG/CONST kernels with Potts or Diagonal
    compatibilities; FULL kernels (the reference forms fg f^T with Eigen's GEMM) and
    Matrix compatibilities within 2e-6 of the largest entry."""
    z = _load("kgrad_small.npz")
    W, H, L, ck, kt1, kt2, nt, n = [int(v) for v in z["c%d_meta" % i]]
    p = "c%d_" % i
    c1, c2 = z[p + "c1"], z[p + "c2"]
    if ck == 0:
        c1, c2 = float(c1[0]), float(c2[0])
    o = po.OracleCRF(W=W, H=H, L=L)
    o.set_unary(z[p + "U"])
    o.add_gaussian(3.0, 2.0, c1, ktype=kt1, ntype=nt)
    o.add_bilateral(20.0, 30.0, 13.0, 17.0, 11.0, z[p + "img"], c2, ktype=kt2, ntype=nt)
    if len(z[p + "kparams"]):
        o.set_kernel_params(z[p + "kparams"])
    np.testing.assert_array_equal(o.kernel_params(), z[p + "kparams"])
    r, ug, cg, kg = o.gradient(n, z[p + "gt"], kernel=True)
    ref = z[p + "kernel_grad"]
    assert kg.shape == ref.shape
    if ck != 2 and 2 not in (kt1, kt2):
        assert r == float(z[p + "obj"])
        assert np.array_equal(bits(ug), bits(z[p + "unary_grad"]))
        assert np.array_equal(bits(kg), bits(ref))
    else:
        assert abs(r - float(z[p + "obj"])) <= 2e-6 * max(1.0, abs(r))
        assert np.abs(kg - ref).max() <= 2e-6 * np.abs(ref).max()
    Q = o.inference(n)
    assert np.abs(Q - z[p + "Q"]).max() <= 2e-6


def test_reference_suite_fixture_10x10():
    """pydensecrf/tests/test_dcrf.py:29-83 on the oracle: the 2-D and the generic paths,
    bit-exact against the reference, and their MAPs agree (test_call_dcrf_eq_dcrf2d)."""
    from pydensecrf_amd import utils
    z = _load("simple_10x10.npz")
    d2 = po.OracleCRF(W=10, H=10, L=2)
    d2.set_unary(z["U"])
    d2.add_bilateral(2.0, 2.0, 2.0, 2.0, 2.0, z["img"], 3.0)
    Q2 = d2.inference(10)
    d = po.OracleCRF(N=100, L=2)
    d.set_unary(z["U"])
    feats = utils.create_pairwise_bilateral(sdims=(2, 2), schan=2, img=z["img"], chdim=2)
    assert np.array_equal(bits(feats), bits(z["feats"]))
    d.add_pairwise(feats, 3.0)
    Qg = d.inference(10)
    assert np.array_equal(bits(Q2), bits(z["Q2d"]))
    assert np.array_equal(bits(Qg), bits(z["Qgeneric"]))
    assert np.array_equal(Q2.argmax(0), Qg.argmax(0))


def _image_check(fname, cfg):
    from pydensecrf_amd import synthetic as syn
    z = _load(fname)
    img, U, _ = syn.make_image(**cfg)
    assert sha(U) == str(z["unary_sha"]) and sha(img) == str(z["image_sha"]), "synthetic generator drifted"
    W, H, L, n = cfg["W"], cfg["H"], cfg["L"], cfg["iters"]
    o = po.OracleCRF(W=W, H=H, L=L)
    o.set_unary(U)
    o.add_gaussian(3.0, 3.0, 3.0)
    o.add_bilateral(80.0, 80.0, 13.0, 13.0, 13.0, img, 10.0)
    lat = o.lattice(0)
    assert lat.V == int(z["V_gaussian"])
    assert sha(lat.offset[:W * H]) == str(z["offset_gaussian_sha"])
    Q = o.inference(n)
    assert np.array_equal(bits(Q[:, z["idx"]]), bits(z["Q_sample"]))
    assert sha(Q.argmax(0).astype(np.int16)) == str(z["map_sha"])


def test_image_320x213_fixture():
    from pydensecrf_amd import synthetic as syn
    _image_check("image_320x213.npz", syn.IMAGE_320x213)


def test_image_1024x768_fixture():
    from pydensecrf_amd import synthetic as syn
    _image_check("image_1024x768.npz", syn.IMAGE_1024x768)


# ---------------------------------------------------------------- live reference
ref = pytest.mark.skipif(not po.ref_available(), reason="oracle/_ref (reference build) not present")


@ref
@pytest.mark.parametrize("d,N", [(2, 203), (3, 117), (5, 333), (7, 61), (10, 45), (15, 30)])
def test_lattice_live_reference(d, N):
    rng = np.random.default_rng(d * 7 + N)
    f = (rng.standard_normal((d, N)) * 4).astype(np.float32)
    a, b = po.OracleLattice(f).arrays, po.RefLattice(f).arrays
    assert a.V == b.V
    assert np.array_equal(a.offset, b.offset) and np.array_equal(a.rank, b.rank)
    assert np.array_equal(bits(a.bary), bits(b.bary)) and np.array_equal(a.nbr, b.nbr)


@ref
@pytest.mark.parametrize("L", [1, 2, 3, 4, 8, 21, 33])
def test_compute_live_reference(L):
    rng = np.random.default_rng(L)
    f = (rng.standard_normal((4, 257)) * 3).astype(np.float32)
    v = rng.standard_normal((L, 257)).astype(np.float32)
    a, b = po.OracleLattice(f), po.RefLattice(f)
    for rev in (False, True):
        assert np.array_equal(bits(a.compute(v, rev)), bits(b.compute(v, rev)))


@ref
def test_generic_model_live_reference():
    rng = np.random.default_rng(5)
    N, L = 500, 4
    U = rng.standard_normal((L, N)).astype(np.float32)
    f1 = (rng.standard_normal((3, N)) * 2).astype(np.float32)
    f2 = (rng.standard_normal((4, N)) * 2).astype(np.float32)
    w = np.array([1, 2, 2, 3], np.float32)
    out = []
    for cls in (po.OracleCRF, po.RefCRF):
        c = cls(N=N, L=L)
        c.set_unary(U)
        c.add_pairwise(f1, w)
        c.add_pairwise(f2, w, ntype=po.NORMALIZE_AFTER)
        out.append(c.inference(7))
    assert np.array_equal(bits(out[0]), bits(out[1]))


@ref
@pytest.mark.parametrize("kt,nt", [(1, 3), (1, 0), (2, 1), (2, 2)])
def test_kernel_gradient_live_reference(kt, nt):
    """Generic features (DenseCRF::addPairwiseEnergy) with set kernel parameters."""
    rng = np.random.default_rng(31 + kt * 4 + nt)
    N, L, n = 400, 4, 3
    U = rng.standard_normal((L, N)).astype(np.float32)
    gt = rng.integers(-1, L, N).astype(np.int16)
    f1 = (rng.standard_normal((3, N)) * 2).astype(np.float32)
    f2 = (rng.standard_normal((4, N)) * 2).astype(np.float32)
    w = np.array([1, 2, 2, 3], np.float32)
    out = []
    for cls in (po.OracleCRF, po.RefCRF):
        c = cls(N=N, L=L) if cls is po.OracleCRF else cls(N=N, L=L, identity_unary=True)
        c.set_unary(U)
        c.add_pairwise(f1, w, ktype=kt, ntype=nt)
        c.add_pairwise(f2, 3.0, ktype=kt, ntype=nt)
        p0 = c.kernel_params()
        c.set_kernel_params(p0 * np.float32(0.9))
        out.append(c.gradient(n, gt, kernel=True))
    (ro, _, _, ko), (rr, _, _, kr) = out
    assert ro == rr
    if kt == 1:
        assert np.array_equal(bits(ko), bits(kr))
    else:
        assert np.abs(ko - kr).max() <= 2e-6 * np.abs(kr).max()


def _learning_case(z, i):
    obj, W, H, L, n = [int(v) for v in z["o%d_meta" % i]]
    p = "o%d_" % i
    o = po.OracleCRF(W=W, H=H, L=L)
    o.set_unary(z[p + "U"])
    o.add_gaussian(3.0, 2.0, 3.0, ntype=3)
    o.add_bilateral(20.0, 30.0, 13.0, 17.0, 11.0, z[p + "img"], 10.0, ntype=3)
    cw = z[p + "class_weight"]
    return o, obj, n, p, (cw if cw.size else None)


@pytest.mark.parametrize("i", range(6))
def test_objective_fixtures_bitexact(i):
    """Hamming and IntersectionOverUnion (objective.cpp:51-108) through DenseCRF::gradient:
    objective, unary, compatibility and kernel gradients bit-exact."""
    z = _load("learning_small.npz")
    o, obj, n, p, cw = _learning_case(z, i)
    r, ug, cg, kg = o.gradient_obj(n, z[p + "gt"], obj, float(z[p + "param"]), cw, kernel=True)
    assert r == float(z[p + "obj"])
    assert np.array_equal(bits(ug), bits(z[p + "unary_grad"]))
    assert np.array_equal(bits(cg), bits(z[p + "compat_grad"]))
    assert np.array_equal(bits(kg), bits(z[p + "kernel_grad"]))


@pytest.mark.parametrize("i", range(2))
def test_logistic_unary_fixture(i):
    """LogisticUnaryEnergy (unary.cpp:48-68): U = L f and its unary gradient sum_t b f^T.
    The oracle takes U as a plain unary (the FP64 product rounded to float; the reference's
    Eigen product rounds in its own order), so Q agrees to 1e-6, and the unary gradient
    is (dL/dU) f^T, the linearity of the sum over iterations, to 1e-5 of its largest
    entry."""
    z = _load("learning_small.npz")
    W, H, L, F, n = [int(v) for v in z["l%d_meta" % i]]
    p = "l%d_" % i
    Lm, f = z[p + "Lm"], z[p + "f"]
    np.testing.assert_array_equal(z[p + "params"], Lm.T.ravel())
    o = po.OracleCRF(W=W, H=H, L=L)
    o.set_unary((Lm.astype(np.float64) @ f.astype(np.float64)).astype(np.float32))
    o.add_gaussian(3.0, 2.0, 3.0, ntype=3)
    o.add_bilateral(20.0, 30.0, 13.0, 17.0, 11.0, z[p + "img"], 10.0, ntype=3)
    assert np.abs(o.inference(n) - z[p + "Q"]).max() <= 1e-6
    r, ug, cg, _ = o.gradient_obj(n, z[p + "gt"], po.LOG_LIKELIHOOD)
    assert abs(r - float(z[p + "obj"])) <= 1e-6 * abs(r)
    g = (ug.astype(np.float64) @ f.T.astype(np.float64)).T.ravel()
    ref = z[p + "unary_param_grad"]
    assert np.abs(g - ref).max() <= 1e-5 * np.abs(ref).max()


def test_debug_energies_fixture():
    """DenseCRF::unaryEnergy / pairwiseEnergy (densecrf.cpp:141-177) from the oracle's
    pairwise messages of the one-hot labelling: bit-exact."""
    z = _load("learning_small.npz")
    W, H, L = [int(v) for v in z["e_meta"]]
    o = po.OracleCRF(W=W, H=H, L=L)
    o.set_unary(z["e_U"])
    o.add_gaussian(3.0, 2.0, 3.0, ntype=3)
    o.add_bilateral(20.0, 30.0, 13.0, 17.0, 11.0, z["e_img"], np.arange(1, L + 1, dtype=np.float32), ntype=1)
    lab = z["e_labels"].astype(np.int64)
    ok = (lab >= 0) & (lab < L)
    idx = np.arange(W * H)
    ue = np.where(ok, z["e_U"][np.clip(lab, 0, L - 1), idx], 0).astype(np.float32)
    assert np.array_equal(bits(ue), bits(z["e_unary"]))
    onehot = np.zeros((L, W * H), np.float32)
    onehot[lab[ok], idx[ok]] = 1
    total = np.zeros(W * H, np.float32)
    for k in range(2):
        m = o.apply(k, onehot)
        e = np.where(ok, np.float32(-0.5) * m[np.clip(lab, 0, L - 1), idx], np.float32(0)).astype(np.float32)
        assert np.array_equal(bits(e), bits(z["e_pair%d" % k]))
        total = (total + e).astype(np.float32)
    assert np.array_equal(bits(total), bits(z["e_pair"]))


def test_gradient_512x512_fixture_reference_vs_fd():
    """The reference's DenseCRF::gradient at 512x512x21 (Matrix compatibilities, term
    weights folded in) against the oracle's FP64 central differences in the same fixture,
    under the bars the GPU is held to: LogLikelihood = -CE, so dCE/dU = -g_U, a pair
    (a, b) of term k moves W_ab = w_k (m_ab + m_ba) / 2 by w_k h, dCE = -w_k g_ab."""
    from pydensecrf_amd import synthetic as syn
    from tests.helpers import fd_relative_errors
    z = _load("grad_512x512.npz")
    _, _, _, c1, c2, kw = syn.gradient_problem_512()
    L = 21
    assert abs(-float(z["ref_obj"]) - float(z["ce_fp64"])) <= 1e-6 * float(z["ce_fp64"])
    npk = L * (L + 1) // 2
    cg = z["ref_compat_grad"].astype(np.float64)
    g_pairs = np.concatenate([-kw[0] * cg[:npk], -kw[1] * cg[npk:]])
    # term weight k: W_k = w_k (m_k + m_k^T) / 2, so dCE/dw_k = -sum_pairs g_ab (m_ab + m_ba) / 2
    iu = np.triu_indices(L)
    g_kw = np.array([-(cg[k * npk:(k + 1) * npk] * (0.5 * (c + c.T))[iu].astype(np.float64)).sum()
                     for k, c in enumerate((c1, c2))])
    rel = fd_relative_errors(-z["ref_unary_grad_sel"].astype(np.float64), g_pairs, g_kw, z)
    assert max(r.max() for r in rel) <= 1e-3, [r.max() for r in rel]


def test_volume_fixture_oracle_bitexact():
    """The oracle reproduces the reference on the volume workload (d = 3 and d = 4
    features, Diagonal compatibilities, 10 iterations): the Q sample bit for bit and
    the MAP hash of the whole volume, for the first volume (the smaller pick); the
    fixture's reference gradient against its FP64 central differences."""
    from pydensecrf_amd import synthetic as syn
    z = _load("volumes_8.npz")
    sizes, fgs, fbs, Us, _ = syn.volume_batch(8)
    i = int(z["picks"][0])
    o = po.OracleCRF(N=sizes[i], L=4)
    o.set_unary(Us[i])
    o.add_pairwise(fgs[i], syn.VOLUME_COMPAT)
    o.add_pairwise(fbs[i], syn.VOLUME_COMPAT)
    Q = o.inference(int(z["iters"]))
    assert np.array_equal(bits(Q[:, z["p0_idx"]]), bits(z["p0_Q_sample"]))
    assert sha(Q.argmax(0).astype(np.int16)) == str(z["p0_map_sha"])
    # the reference's compatibility gradient (float sums over ~5e5 voxels) against the
    # FP64 central differences in the same fixture: within 3e-4 of the largest entry
    for p in ("p0_", "p1_"):
        fd = z[p + "compat_fd"]
        assert np.abs(-z[p + "compat_grad"].astype(np.float64) - fd).max() <= 3e-4 * np.abs(fd).max()
