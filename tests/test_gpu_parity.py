"""GPU parity: the HIP path (through the C ABI) against the oracle on the same inputs.

Bars: the lattice (vertex offsets, ranks, barycentric weights, blur neighbours, vertex
keys, V) is bit-exact; Q, step temporaries, pairwise messages and KL within the
tolerances stated per test.  The oracle is bit-exact with the reference's own C++
(tests/test_oracle_reference.py), so these are parity checks against the reference.
"""
import numpy as np
import pytest

from oracle import pyoracle as po
from tests.helpers import assert_lattice_equal, build_both, image_case

pytestmark = pytest.mark.gpu

SYM, BEFORE, AFTER, NONE = 3, 1, 2, 0
Q_TOL = 1e-5  # absolute, on Q (probabilities)


def _std_terms(img, nt=SYM, cg=3.0, cb=10.0):
    return [("gauss", 3.0, 3.0, cg, nt), ("bilat", 80.0, 80.0, 13.0, 13.0, 13.0, img, cb, nt)]


LATTICE_CASES = [
    # (W, H, L, description)
    (10, 10, 2, "reference test size"),
    (13, 7, 3, "N % 4 == 3: pad lanes"),
    (31, 17, 5, "N % 4 == 3"),
    (40, 30, 21, "small"),
    (320, 213, 21, "C0 size"),
]


@pytest.mark.parametrize("W,H,L,desc", LATTICE_CASES)
def test_lattice_bitexact_2d(W, H, L, desc):
    img, U, _ = image_case(W, H, L, seed=W * 31 + H)
    g, o = build_both(W, H, L, U, _std_terms(img))
    for k in range(2):
        assert_lattice_equal(g.lattice(k), o.lattice(k), W * H)


@pytest.mark.parametrize("d,N,scale", [(1, 37, 2.0), (3, 150, 1.5), (4, 401, 1.0), (6, 255, 0.7), (9, 97, 0.5),
                                       (15, 64, 0.4)])
def test_lattice_bitexact_generic(d, N, scale):
    rng = np.random.default_rng(d * 1000 + N)
    f = (rng.standard_normal((d, N)) * scale * 4).astype(np.float32)
    L = 3
    U = rng.standard_normal((L, N)).astype(np.float32)
    g, o = build_both(N, 1, L, U, [("feat", f, 1.0, SYM)], generic=True)
    assert_lattice_equal(g.lattice(0), o.lattice(0), N)


def test_norm_matches_oracle():
    W, H, L = 64, 48, 4
    img, U, _ = image_case(W, H, L, seed=7)
    for nt in (SYM, BEFORE, AFTER, NONE):
        g, o = build_both(W, H, L, U, _std_terms(img, nt))
        for k in range(2):
            gn = g.lattice(k)["norm"]
            on = o.norm(k)
            np.testing.assert_allclose(gn, on, rtol=2e-6, atol=0, err_msg="ntype %d term %d" % (nt, k))


@pytest.mark.parametrize("W,H,L,n", [(10, 10, 2, 10), (13, 7, 3, 5), (40, 30, 21, 5), (64, 48, 4, 8)])
@pytest.mark.parametrize("nt", [SYM, BEFORE, AFTER, NONE])
def test_inference_potts(W, H, L, n, nt):
    img, U, _ = image_case(W, H, L, seed=W + H + nt)
    g, o = build_both(W, H, L, U, _std_terms(img, nt))
    Qg = g.inference(n)
    Qo = o.inference(n)
    assert np.abs(Qg - Qo).max() <= Q_TOL


@pytest.mark.parametrize("L", [1, 2, 3, 5, 21])
@pytest.mark.parametrize("ck", ["diag", "matrix"])
def test_inference_compat_kinds(L, ck):
    W, H = 37, 23
    img, U, _ = image_case(W, H, L, seed=L)
    rng = np.random.default_rng(L + 17)
    if ck == "diag":
        c1, c2 = rng.uniform(0.5, 3, L).astype(np.float32), rng.uniform(1, 6, L).astype(np.float32)
    else:
        c1 = (-np.eye(L) * 2 + rng.normal(0, 0.3, (L, L))).astype(np.float32)
        c2 = (-np.eye(L) * 5 + rng.normal(0, 0.5, (L, L))).astype(np.float32)
    g, o = build_both(W, H, L, U, [("gauss", 3.0, 3.0, c1, SYM), ("bilat", 20.0, 20.0, 13., 13., 13., img, c2, SYM)])
    assert np.abs(g.inference(6) - o.inference(6)).max() <= Q_TOL


def test_inference_c0_size():
    from pydensecrf_amd import synthetic as syn
    img, U, _ = syn.make_image(**syn.IMAGE_320x213)
    g, o = build_both(320, 213, 21, U, _std_terms(img))
    Qg, Qo = g.inference(5), o.inference(5)
    err = np.abs(Qg - Qo).max()
    print("C0 max |Q_gpu - Q_ref| = %.3g" % err)
    assert err <= Q_TOL
    assert (Qg.argmax(0) != Qo.argmax(0)).sum() == 0


def test_generic_features_inference():
    rng = np.random.default_rng(3)
    Z, Y, X = 6, 9, 11
    N, L = Z * Y * X, 4
    zz, yy, xx = np.meshgrid(np.arange(Z), np.arange(Y), np.arange(X), indexing="ij")
    inten = rng.uniform(0, 1, (Z, Y, X)).astype(np.float32)
    fg = (np.stack([zz, yy, xx]).reshape(3, -1) / 3.0).astype(np.float32)
    fb = np.concatenate([np.stack([zz, yy, xx]).reshape(3, -1) / 4.0, inten.reshape(1, -1) / 0.1]).astype(np.float32)
    U = rng.standard_normal((L, N)).astype(np.float32)
    w = np.array([1, 2, 2, 3], np.float32)
    g, o = build_both(N, 1, L, U, [("feat", fg, w, SYM), ("feat", fb, w, SYM)], generic=True)
    assert np.abs(g.inference(10) - o.inference(10)).max() <= Q_TOL


def test_start_step_kl_apply():
    W, H, L = 41, 29, 6
    img, U, _ = image_case(W, H, L, seed=11)
    g, o = build_both(W, H, L, U, _std_terms(img))
    Qg, Qo = g.start(), o.start()
    assert np.abs(Qg - Qo).max() <= 1e-6
    for _ in range(3):
        Qg, t1g, t2g = g.step(Qg)
        Qo, t1o, t2o = o.step(Qo)
        assert np.abs(Qg - Qo).max() <= Q_TOL
        np.testing.assert_allclose(t1g, t1o, rtol=1e-5, atol=1e-4)
        np.testing.assert_allclose(t2g, t2o, rtol=1e-5, atol=1e-4)
        klg, klo = g.kl(Qg), o.kl(Qo)
        assert abs(klg - klo) <= 1e-5 * max(1.0, abs(klo)), (klg, klo)
    for k in range(2):
        for tr in (False, True):
            np.testing.assert_allclose(g.apply(k, Qo, tr), o.apply(k, Qo, tr), rtol=1e-5, atol=1e-5)


@pytest.mark.parametrize("W,H,L", [(512, 512, 21), (1001, 701, 13), (2048, 1600, 21)])
def test_large_host_transfers_bitwise(W, H, L):
    """Host arrays of 22 MB (512x512x21), 36.5 MB (1001x701x13, not a multiple of the
    chunks' 4 KiB alignment) and 275 MB (2048x1600x21: two windows of the 256 MiB staging
    area), above the 8 MiB from which the library moves them through its page-locked
    staging area in chunks with several host threads: the unary upload,
    inference's download, start / step (Q up, Q, tmp1 and tmp2 down), KL and apply (Q
    up, the messages down).  Each agrees bit for bit with a path whose arrays never
    leave the device or with the other host path, and the staged unary gives the same Q
    as the unary passed as a device buffer."""
    import torch

    from pydensecrf_amd import capi

    N = W * H
    img, U, _ = image_case(W, H, L, seed=23)
    g = capi.Model.create_2d(W, H, L)
    g.set_unary(U)
    g.add_gaussian(3.0, 3.0, 3.0)
    g.add_bilateral(80.0, 80.0, 13.0, 13.0, 13.0, img, 10.0)
    Qi = g.inference(2)
    assert Qi.shape == (L, N)
    Qd = torch.empty((N, L), dtype=torch.float32, device="cuda")
    Ud = torch.tensor(np.ascontiguousarray(U.T), device="cuda")
    g.inference_device(2, Qd.data_ptr())
    torch.cuda.synchronize()
    assert np.array_equal(Qi.T.view(np.int32), Qd.cpu().numpy().view(np.int32))
    g.inference_device(2, Qd.data_ptr(), Ud.data_ptr())  # the unary from the device
    torch.cuda.synchronize()
    assert np.array_equal(Qi.T.view(np.int32), Qd.cpu().numpy().view(np.int32))
    Q = g.start()
    for _ in range(2):
        Q, t1, t2 = g.step(Q)
    assert np.array_equal(Qi.view(np.int32), Q.view(np.int32))
    assert np.isfinite(t1).all() and np.isfinite(t2).all() and np.abs(t2).max() > 0
    kl1, kl2 = g.kl(Q), g.kl(np.ascontiguousarray(Q))
    assert kl1 == kl2 and np.isfinite(kl1)
    for k in range(2):
        a = g.apply(k, Q)
        assert np.array_equal(a.view(np.int32), g.apply(k, np.ascontiguousarray(Q)).view(np.int32))
        assert np.abs(a).max() > 0


def test_map_and_current_map():
    """DenseCRF::map / currentMap (densecrf.cpp:132-137, 202-211): the first label of each
    pixel's largest Q.  Exact ties take the first label (Eigen's maxCoeff); one label
    gives zeros; a batched model's labels are its instances' in order; the device form
    gives the host form's labels."""
    import torch

    from pydensecrf_amd import capi

    W, H, L, n = 45, 31, 6, 3
    img, U, _ = image_case(W, H, L, seed=13)
    g, o = build_both(W, H, L, U, _std_terms(img))
    Q = g.inference(n)
    assert np.array_equal(g.map(n), Q.argmax(0).astype(np.int16))
    assert np.array_equal(g.map(n), o.inference(n).argmax(0).astype(np.int16))
    # ties and equal rows: the first of the largest
    N = W * H
    rng = np.random.default_rng(5)
    T = rng.integers(0, 3, size=(L, N)).astype(np.float32)  # many exact ties
    T[:, :7] = 1.0  # whole columns equal
    want = np.array([next(l for l in range(L) if T[l, p] == T[:, p].max()) for p in range(N)], np.int16)
    assert np.array_equal(g.current_map(T), want) and np.array_equal(g.current_map(T)[:7], np.zeros(7, np.int16))
    Td = torch.tensor(np.ascontiguousarray(T.T), device="cuda")
    lab = torch.empty(N, dtype=torch.int16, device="cuda")
    g.current_map_device(Td.data_ptr(), lab.data_ptr())
    torch.cuda.synchronize()
    assert np.array_equal(lab.cpu().numpy(), want)
    # one label: every pixel takes label 0
    g1, _ = build_both(W, H, 1, U[:1], _std_terms(img))
    assert not g1.map(2).any()
    # a batch: labels of the instances in order
    from pydensecrf_amd.batched import BatchedDenseCRF2D
    shapes = [(20, 15), (33, 9)]
    cases = [image_case(w, h, L, seed=60 + i) for i, (w, h) in enumerate(shapes)]
    b = BatchedDenseCRF2D(shapes, L)
    b.setUnaryEnergy([c[1] for c in cases])
    b.addPairwiseGaussian(sxy=3, compat=3)
    b.addPairwiseBilateral(sxy=30, srgb=13, rgbims=[c[0] for c in cases], compat=10)
    Qb = b.inference(n)
    assert np.array_equal(b.model.map(n), np.concatenate([q.argmax(0) for q in Qb]).astype(np.int16))


def test_release_cached_memory_keeps_models_working():
    """dcrf_release_cached_memory returns the cached device blocks and the page-locked
    staging area (which a 22 MB host transfer has just grown); a live model then gives
    the same bits again, its staged transfers taking a new staging area."""
    import ctypes

    from pydensecrf_amd import capi

    W, H, L = 512, 512, 21
    img, U, _ = image_case(W, H, L, seed=31)
    g = capi.Model.create_2d(W, H, L)
    g.set_unary(U)
    g.add_gaussian(3.0, 3.0, 3.0)
    g.add_bilateral(80.0, 80.0, 13.0, 13.0, 13.0, img, 10.0)
    Q0 = g.inference(2).copy()
    freed = ctypes.c_size_t(0)
    capi.check(capi.lib().dcrf_release_cached_memory(ctypes.byref(freed)))
    assert freed.value >= W * H * L * 4  # at least the staging area of Q's download
    g.set_unary(U)  # staged again at the next computation
    assert np.array_equal(g.inference(2).view(np.int32), Q0.view(np.int32))
    capi.check(capi.lib().dcrf_release_cached_memory(None))


@pytest.mark.parametrize("terms", ["g", "b", "gb", "bg"])
@pytest.mark.parametrize("L", [4, 21])
@pytest.mark.parametrize("W,H", [(45, 31), (64, 48), (37, 23)])
def test_fused_meanfield_bitwise(terms, L, W, H):
    """inference() runs the fused mean-field kernel for these term layouts, step() the
    generic one: the same arithmetic in the same order, so Q agrees bit for bit; so do
    the pixel-run form (whole and partial runs, two run heights) and the generic kernel
    when selected."""
    img, U, _ = image_case(W, H, L, seed=len(terms) * 7 + L)
    spec = {"g": ("gauss", 3.0, 3.0, 3.0, SYM), "b": ("bilat", 20.0, 20.0, 13.0, 13.0, 13.0, img, 10.0, SYM)}
    g, _ = build_both(W, H, L, U, [spec[c] for c in terms])
    assert g.splat_layout()["meanfield"] == 1  # the fused form by default
    Qi = g.inference(3)
    Q = g.start()
    for _ in range(3):
        Q, _, _ = g.step(Q)
    assert np.array_equal(Qi.view(np.int32), Q.view(np.int32))
    for mode, rows, path in ((3, 0, 2), (3, 4, 2), (1, 0, 0)):  # pixel runs, generic
        g.set_meanfield_kernel(mode, rows)
        assert g.splat_layout()["meanfield"] == path
        assert np.array_equal(g.inference(3).view(np.int32), Qi.view(np.int32)), (mode, rows)


def test_fused_meanfield_generic_runs_bitwise():
    """A generic (not 2-D) model whose one term has d = 2: the pixel-run kernel runs
    over runs of consecutive pixels (a partial last run included)."""
    rng = np.random.default_rng(29)
    Wg, Hg, L = 50, 41, 21
    yy, xx = np.mgrid[0:Hg, 0:Wg]
    f = (np.stack([xx, yy]).reshape(2, -1) / 3.0).astype(np.float32)
    U = rng.standard_normal((L, Wg * Hg)).astype(np.float32)
    g, o = build_both(Wg * Hg, 1, L, U, [("feat", f, 3.0, SYM)], generic=True)
    g.set_meanfield_kernel(3)
    assert g.splat_layout()["meanfield"] == 2
    Qi = g.inference(3)
    Q = g.start()
    for _ in range(3):
        Q, _, _ = g.step(Q)
    assert np.array_equal(Qi.view(np.int32), Q.view(np.int32))
    assert np.abs(Qi - o.inference(3)).max() <= Q_TOL


def test_fused_meanfield_without_runs_bitwise():
    """Random 5-D features share almost no lattice vertices between neighbouring
    pixels, so even when pixel runs are requested a run's rows exceed the LDS budget:
    inference takes the plain fused kernel (interleaved slice pairs, rows gathered from
    memory), bit-identical to the generic kernel of step()."""
    rng = np.random.default_rng(23)
    N, L = 3000, 21
    f = (rng.standard_normal((5, N)) * 6).astype(np.float32)
    U = rng.standard_normal((L, N)).astype(np.float32)
    g, o = build_both(N, 1, L, U, [("feat", f, 5.0, SYM)], generic=True)
    g.set_meanfield_kernel(3)
    assert g.splat_layout()["meanfield"] == 1  # the runs' rows exceed the LDS budget
    Qi = g.inference(3)
    Q = g.start()
    for _ in range(3):
        Q, _, _ = g.step(Q)
    assert np.array_equal(Qi.view(np.int32), Q.view(np.int32))
    assert np.abs(Qi - o.inference(3)).max() <= Q_TOL


def test_no_unary_and_no_terms():
    W, H, L = 9, 7, 3
    g, o = build_both(W, H, L, None, [])
    np.testing.assert_allclose(g.inference(3), o.inference(3), atol=1e-7)
    assert np.all(g.start() == 0)


def test_deterministic():
    W, H, L = 96, 80, 21
    img, U, _ = image_case(W, H, L, seed=5)
    g, _ = build_both(W, H, L, U, _std_terms(img))
    a = g.inference(5)
    b = g.inference(5)
    assert np.array_equal(a.view(np.int32), b.view(np.int32))


@pytest.mark.parametrize("chunk", [1, 7, 64, 100000])
def test_splat_chunk_invariance(chunk):
    W, H, L = 120, 90, 21
    img, U, _ = image_case(W, H, L, seed=9)
    g, o = build_both(W, H, L, U, _std_terms(img))
    g.set_splat_chunk(chunk)
    assert np.abs(g.inference(5) - o.inference(5)).max() <= Q_TOL


@pytest.mark.parametrize("L", [21, 4, 1, 64, 300])
def test_heavy_vertex_recombination(L):
    """Splat segments of one contributor each (chunk 1): every vertex with more than 16
    contributors is recombined from more than 16 partial rows, in one launch with a
    whole wave per such vertex (wave-mode shapes: L = 21, 4, 1, 64) or in two levels
    (L = 300, groups wider than a wave).  Q and the transposed message against the
    oracle."""
    W, H = 64, 48
    img, U, _ = image_case(W, H, L, seed=L + 300)
    g, o = build_both(W, H, L, U, _std_terms(img))
    g.set_splat_chunk(1)
    assert np.abs(g.inference(4) - o.inference(4)).max() <= Q_TOL
    Q = o.start()
    np.testing.assert_allclose(g.apply(1, Q, True), o.apply(1, Q, True), rtol=1e-5, atol=1e-5)


@pytest.mark.parametrize("L", [4, 7, 19, 64, 65, 130, 300])
@pytest.mark.parametrize("ck", ["potts", "matrix"])
def test_label_count_shapes(L, ck):
    """Labels per lane and group layouts: L = 4 (one lane per pixel), 7 / 19 (a partial
    last lane), 64, 65, 130 and 300 (groups wider than a wave, LDS path)."""
    W, H = 23, 17
    img, U, _ = image_case(W, H, L, seed=L + 100)
    if ck == "potts":
        c1, c2 = 3.0, 10.0
    else:
        rng = np.random.default_rng(L)
        c1 = (-2 * np.eye(L) + rng.normal(0, 0.05, (L, L))).astype(np.float32)
        c2 = (-4 * np.eye(L) + rng.normal(0, 0.05, (L, L))).astype(np.float32)
    g, o = build_both(W, H, L, U, [("gauss", 3.0, 3.0, c1, SYM), ("bilat", 20.0, 20.0, 13., 13., 13., img, c2, SYM)])
    assert np.abs(g.inference(4) - o.inference(4)).max() <= Q_TOL
    Q = o.start()
    assert abs(g.kl(Q) - o.kl(Q)) <= 1e-5 * max(1.0, abs(o.kl(Q)))
    np.testing.assert_allclose(g.apply(1, Q, True), o.apply(1, Q, True), rtol=1e-5, atol=1e-5)


@pytest.mark.parametrize("tile", [0, 1, 5, 16, 24, 32, 4096])
@pytest.mark.parametrize("generic", [False, True])
def test_splat_tile_invariance(tile, generic):
    """Splat segments cut at tile boundaries give the same messages up to summation
    order: sides 1 (a segment per contributor), 5 (256-thread workgroups) and 16
    (1024-thread workgroups) run the LDS-tiled splat; 0 (the default, 128) and 4096 (one
    tile per image) exceed its LDS and run the global form; 24 and 32 run either,
    depending on the workgroup LDS limit and the term layout."""
    W, H, L = 120, 90, 21
    img, U, _ = image_case(W, H, L, seed=19)
    if generic:
        yy, xx = np.mgrid[0:H, 0:W]
        f = np.concatenate([np.stack([xx, yy]).reshape(2, -1) / 20.0,
                            img.reshape(-1, 3).T / 13.0]).astype(np.float32)
        terms = [("feat", f, 10.0, SYM)]
    else:
        terms = _std_terms(img)
    g, o = build_both(W, H, L, U, terms, generic=generic, tile=tile)
    g.set_splat_chunk(16)
    lay = g.splat_layout()
    if tile in (1, 5, 16, 0, 4096):
        assert lay["tiled"] == (tile in (1, 5, 16)), lay
    assert np.abs(g.inference(5) - o.inference(5)).max() <= Q_TOL
    np.testing.assert_allclose(g.apply(0, o.start(), True), o.apply(0, o.start(), True), rtol=1e-5, atol=1e-5)


def test_autotuned_chunk_keeps_results(tmp_path, monkeypatch):
    from pydensecrf_amd import tune
    monkeypatch.setenv("PYDENSECRF_AMD_TUNE_CACHE", str(tmp_path / "tune.json"))
    W, H, L = 96, 72, 21
    img, U, _ = image_case(W, H, L, seed=77)
    g, o = build_both(W, H, L, U, _std_terms(img))
    best, times = tune.tune_splat_chunk(g, iters=2, reps=2)
    assert best in tune.CHUNKS and set(times) == set(tune.CHUNKS)
    assert np.abs(g.inference(5) - o.inference(5)).max() <= Q_TOL
    again, t2 = tune.tune_splat_chunk(g)
    assert again == best and t2 is None  # cached


@pytest.mark.parametrize("dims,scales", [((2, 5), (1.5, 1.5)), ((1, 3), (1.5, 1.5)), ((4, 6), (1.5, 1.5)),
                                         ((7,), (1.5,)), ((2, 5), (12.0, 0.3)), ((2, 3, 5), (12.0, 1.0, 0.3))])
def test_blur_pairs_bitwise(dims, scales):
    """Multi-step blur passes (at most two steps, the default; at most three) against
    one pass per step: the same bits for inference, the transposed and forward messages,
    and the gradient (which runs the transposed filter), for feature dimensions with an
    odd and an even number of steps, mixed in one model, and lattices of very different
    vertex counts (scales 12 and 0.3)."""
    from pydensecrf_amd import capi
    rng = np.random.default_rng(sum(dims))
    N, L = 3000, 5
    U = rng.standard_normal((L, N)).astype(np.float32)
    gt = rng.integers(0, L, N).astype(np.int16)
    feats = [(rng.standard_normal((d, N)) * sc).astype(np.float32) for d, sc in zip(dims, scales)]
    out = []
    for pairs, steps in ((False, 2), (True, 2), (True, 3)):
        g = capi.Model.create(N, L)
        g.set_unary(U)
        for i, f in enumerate(feats):
            g.add_pairwise(f, 2.0 + i, ntype=[SYM, AFTER, BEFORE, NONE][i % 4])
        g.set_blur_pairs(pairs)
        g.set_blur_steps(steps)
        Q = g.inference(4)
        msgs = [g.apply(k, Q, tr) for k in range(len(feats)) for tr in (False, True)]
        grads = g.gradient(3, gt)
        out.append((Q, msgs, grads))
    Qa, ma, ga = out[0]
    for Qb, mb, gb in out[1:]:
        np.testing.assert_array_equal(Qa, Qb)
        for x, y in zip(ma, mb):
            np.testing.assert_array_equal(x, y)
        assert ga[0] == gb[0]
        for x, y in zip(ga[1:], gb[1:]):
            np.testing.assert_array_equal(x, y)


@pytest.mark.parametrize("W,H,L,n", [(203, 149, 21, 5), (64, 48, 4, 3), (9, 7, 5, 4), (1024, 768, 21, 10)])
def test_q_blocks_bitwise(W, H, L, n):
    """Q held in 8 x 8 pixel blocks between iterations (dcrf_set_q_blocks, image sides
    not multiples of 8 included): the same bits as the default layout, through the host
    path, the device path and a captured graph."""
    import torch
    from pydensecrf_amd import capi
    img, U, _ = image_case(W, H, L, seed=W + H)
    out = []
    for on in (False, True):
        g = capi.Model.create_2d(W, H, L)
        g.set_unary(U)
        g.add_gaussian(3.0, 3.0, 3.0)
        g.add_bilateral(80.0, 80.0, 13.0, 13.0, 13.0, img, 10.0)
        g.set_q_blocks(on)
        Qh = np.ascontiguousarray(g.inference(n))
        Qd = torch.full((W * H * L,), -1.0, device="cuda")
        g.inference_device(n, Qd.data_ptr())
        Qg = torch.full_like(Qd, -1.0)
        g.capture(n, Qg.data_ptr())
        g.launch_graph()
        torch.cuda.synchronize()
        out.append((Qh, Qd.cpu().numpy(), Qg.cpu().numpy()))
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(np.asarray(a).view(np.int32).ravel(), np.asarray(b).view(np.int32).ravel())
    assert np.array_equal(out[1][0].T.ravel().view(np.int32), out[1][1].view(np.int32))


def test_autotuned_knobs_keep_results(tmp_path, monkeypatch):
    """The widened tuner: the bit-identical knobs (blur steps per pass, mean-field kernel,
    Q in blocks) tuned on a model leave Q's bits as they were; the cached winners are
    applied by a new model of the same shape before its first inference, and the
    build knobs (splat tile side) by a new model of the same size at creation."""
    from pydensecrf_amd import capi, tune
    monkeypatch.setenv("PYDENSECRF_AMD_TUNE_CACHE", str(tmp_path / "tune.json"))
    monkeypatch.setenv("PYDENSECRF_AMD_TUNE", "1")
    W, H, L = 96, 72, 21
    img, U, _ = image_case(W, H, L, seed=78)
    g, o = build_both(W, H, L, U, _std_terms(img))
    ref = np.ascontiguousarray(g.inference(5))
    knobs = ("blur_pairs", "blur_steps", "meanfield", "q_blocks")
    best, times = tune.tune_model(g, iters=2, reps=2, knobs=knobs)
    assert set(best) == set(knobs) and all(set(times[k]) == set(tune.MODEL_KNOBS[k]) for k in knobs)
    assert np.array_equal(np.ascontiguousarray(g.inference(5)).view(np.int32), ref.view(np.int32))
    again, t2 = tune.tune_model(g, knobs=knobs)
    assert again == best and t2 is None  # cached
    g2, _ = build_both(W, H, L, U, _std_terms(img))
    Q2 = np.ascontiguousarray(g2.inference(5))  # applies the cached knobs first
    assert g2.splat_layout()["meanfield"] == (2 if best["meanfield"] == 3 else 1)
    assert np.array_equal(Q2.view(np.int32), ref.view(np.int32))

    def build(side):
        m = capi.Model.create_2d(W, H, L)
        m.set_splat_tile(side)
        m.set_unary(U)
        m.add_gaussian(3.0, 3.0, 3.0)
        m.add_bilateral(80.0, 80.0, 13.0, 13.0, 13.0, img, 10.0)
        return m
    side, ts = tune.tune_splat_tile(build, candidates=(64, 128), iters=2, reps=2)
    assert set(ts) == {64, 128}
    g3, _ = build_both(W, H, L, U, _std_terms(img))
    assert g3.splat_layout()["tile_side"] == side
    assert np.abs(g3.inference(5) - o.inference(5)).max() <= Q_TOL


@pytest.mark.parametrize("R,below", [(1, 0), (1, 50000), (3, 0)])
def test_blur_lanes_bitwise(R, below):
    """Blur passes with another lane split (dcrf_set_blur_lanes): the blur is elementwise
    per label, so inference, the gradient (transposed filter) and a captured graph give
    the same bits as the default split."""
    import torch
    from pydensecrf_amd import capi
    W, H, L = 130, 97, 21
    img, U, gt = image_case(W, H, L, seed=123)
    out = []
    for knob in (None, (R, below)):
        g = capi.Model.create_2d(W, H, L)
        g.set_unary(U)
        g.add_gaussian(3.0, 3.0, 3.0)
        g.add_bilateral(80.0, 80.0, 13.0, 13.0, 13.0, img, 10.0)
        if knob:
            g.set_blur_lanes(*knob)
        Q = np.ascontiguousarray(g.inference(5))
        obj, ug, cg, kg = g.gradient(3, gt)
        Qg = torch.full((W * H * L,), -1.0, device="cuda")
        g.capture(5, Qg.data_ptr())
        g.launch_graph()
        torch.cuda.synchronize()
        out.append((Q, ug, cg, Qg.cpu().numpy(), obj))
    for a, b in zip(out[0][:4], out[1][:4]):
        assert np.array_equal(np.asarray(a).view(np.int32), np.asarray(b).view(np.int32))
    assert out[0][4] == out[1][4]
    with pytest.raises(ValueError):
        g.set_blur_lanes(2)  # 2 does not divide 21
