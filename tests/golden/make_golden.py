"""Writes the golden fixtures of this directory from the reference's own C++.

The reference mean-field code (lucasb-eyer/pydensecrf, pydensecrf/densecrf/src) is
compiled from its sources by oracle/Makefile (`make -C oracle ref`) into
oracle/_ref/libdcrf_ref.so; this script drives that build through
oracle/pyoracle.py (RefLattice / RefCRF) and stores inputs and outputs as .npz data.

    make -C oracle ref && python tests/golden/make_golden.py [fixture ...]

Fixtures
  lattice_small.npz  Permutohedral::init on small feature sets (d = 1..6, N % 4 != 0
                     included): offset, rank, barycentric, blur neighbours, V.
  crf_small.npz      DenseCRF2D / DenseCRF cases (Potts / Diagonal / Matrix, all
                     normalisations, L = 1, 2, 3, 5, 21): Q after n iterations, one
                     stepInference (Q, tmp1, tmp2), KL, DenseCRF::gradient (LogLikelihood,
                     identity-gradient unary) for unary and compatibility parameters.
  kgrad_small.npz    Permutohedral::gradient on small feature sets, and
                     DenseCRF::gradient's kernel_grad (DIAG / FULL / CONST kernels, all
                     normalisations, default and set kernel parameters) with the
                     objective and unary / compatibility gradients of the same call.
  learning_small.npz DenseCRF::gradient with the Hamming (counted and given class
                     weights) and IntersectionOverUnion objectives; a LogisticUnaryEnergy
                     model (its parameters, Q and its unary gradient sum_t b f^T);
                     unaryEnergy / pairwiseEnergy of a labelling.
  simple_10x10.npz   the reference test suite's own fixture (test_dcrf.py:8-26):
                     unary, image and Q of test_call_dcrf2d / test_call_dcrf.
  image_320x213.npz  the 320x213x21 5-iteration workload: V per term, SHA-256 of the
                     lattice offsets, Q on a seeded sample of 4096 pixels.
  image_1024x768.npz the 1024x768x21 10-iteration workload: V per term, Q sample.
  batch_64.npz       the 64-image batch: Q samples and MAP hashes of three instances.
  volumes_8.npz      the 8-volume batch: Q samples, MAP hashes and DenseCRF::gradient
                     of two volumes.
  grad_512x512.npz   the 512x512x21 Matrix-compatibility workload: DenseCRF::gradient,
                     and FP64 central differences of the cross-entropy from the oracle.
  lbfgs_small.npz    minimizeLBFGS (liblbfgs 1.10): evaluated points and results on
                     Rosenbrock and on a small CRF learning energy.
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import pyoracle as po  # noqa: E402
from pydensecrf_amd import synthetic as syn  # noqa: E402
from pydensecrf_amd import utils  # noqa: E402


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def lattice_small():
    out = {}
    rng = np.random.default_rng(20240)
    cases = [(1, 9, 3.0), (2, 37, 2.0), (3, 64, 1.5), (4, 50, 1.0), (5, 101, 2.0), (6, 23, 0.8)]
    for i, (d, N, s) in enumerate(cases):
        f = (rng.standard_normal((d, N)) * s * 3).astype(np.float32)
        lat = po.RefLattice(f).arrays
        out["c%d_feat" % i] = f
        out["c%d_offset" % i] = lat.offset
        out["c%d_rank" % i] = lat.rank
        out["c%d_bary" % i] = lat.bary
        out["c%d_nbr" % i] = lat.nbr
        out["c%d_V" % i] = np.int32(lat.V)
    out["ncases"] = np.int32(len(cases))
    np.savez_compressed(os.path.join(HERE, "lattice_small.npz"), **out)


CRF_CASES = [
    # (W, H, L, compat kind, ntype, n_iter)
    (10, 10, 2, "potts", 3, 10),
    (13, 7, 3, "potts", 3, 5),
    (13, 7, 3, "diag", 1, 5),
    (17, 9, 5, "matrix", 2, 4),
    (17, 9, 1, "potts", 0, 3),
    (24, 20, 21, "potts", 3, 5),
    (24, 20, 21, "diag", 3, 5),
    (24, 20, 21, "matrix", 3, 3),
]


def crf_small():
    out = {}
    for i, (W, H, L, ck, nt, n) in enumerate(CRF_CASES):
        rng = np.random.default_rng(100 + i)
        img, U, gt = syn.make_image(W, H, L=L, sites=4, seed=100 + i)
        if ck == "potts":
            c1, c2 = np.float32(3.0), np.float32(10.0)
        elif ck == "diag":
            c1 = rng.uniform(0.5, 3, L).astype(np.float32)
            c2 = rng.uniform(1, 8, L).astype(np.float32)
        else:
            c1 = (-2 * np.eye(L) + rng.normal(0, 0.3, (L, L))).astype(np.float32)
            c2 = (-5 * np.eye(L) + rng.normal(0, 0.5, (L, L))).astype(np.float32)
        crf = po.RefCRF(W=W, H=H, L=L, identity_unary=True)
        crf.set_unary(U)
        cc1 = float(c1) if ck == "potts" else c1
        cc2 = float(c2) if ck == "potts" else c2
        crf.add_gaussian(3.0, 2.0, cc1, ntype=nt)
        crf.add_bilateral(20.0, 30.0, 13.0, 17.0, 11.0, img, cc2, ntype=nt)
        Q = crf.inference(n)
        Qs = crf.start()
        Q1, t1, t2 = crf.step(Qs)
        kl = crf.kl(Q)
        r, ug, cg = crf.gradient(n, gt)
        p = "c%d_" % i
        out.update({p + "img": img, p + "U": U, p + "gt": gt, p + "c1": np.atleast_1d(c1), p + "c2": np.atleast_1d(c2),
                    p + "Q": Q, p + "Q1": Q1, p + "tmp1": t1, p + "tmp2": t2, p + "kl": np.float64(kl),
                    p + "obj": np.float64(r), p + "unary_grad": ug, p + "compat_grad": cg,
                    p + "meta": np.array([W, H, L, ["potts", "diag", "matrix"].index(ck), nt, n], np.int32)})
    out["ncases"] = np.int32(len(CRF_CASES))
    np.savez_compressed(os.path.join(HERE, "crf_small.npz"), **out)


KG_CASES = [
    # (W, H, L, compat kind, ktype gaussian, ktype bilateral, ntype, n_iter, kernel-parameter mode)
    (11, 9, 3, "potts", 1, 1, 3, 3, "default"),
    (11, 9, 3, "diag", 1, 1, 1, 3, "scaled"),
    (13, 8, 4, "matrix", 2, 2, 2, 2, "scaled"),
    (9, 9, 2, "potts", 2, 2, 0, 3, "default"),
    (16, 12, 5, "matrix", 1, 1, 3, 4, "scaled"),
    (10, 10, 3, "potts", 0, 1, 1, 2, "scaled"),
    (14, 10, 21, "potts", 1, 2, 3, 2, "scaled"),
    (12, 7, 1, "potts", 1, 1, 2, 3, "default"),
]


def kgrad_small():
    out = {}
    rng = np.random.default_rng(4321)
    lat_cases = [(2, 61, 1), (3, 80, 2), (5, 97, 4), (4, 40, 3)]
    for i, (d, N, L) in enumerate(lat_cases):
        f = (rng.standard_normal((d, N)) * 3).astype(np.float32)
        a = rng.standard_normal((L, N)).astype(np.float32)
        b = rng.standard_normal((L, N)).astype(np.float32)
        out["l%d_feat" % i], out["l%d_a" % i], out["l%d_b" % i] = f, a, b
        out["l%d_df" % i] = po.RefLattice(f).gradient(a, b)
    out["nlat"] = np.int32(len(lat_cases))
    for i, (W, H, L, ck, kt1, kt2, nt, n, mode) in enumerate(KG_CASES):
        r2 = np.random.default_rng(700 + i)
        img, U, gt = syn.make_image(W, H, L=L, sites=4, seed=700 + i)
        if ck == "potts":
            c1, c2 = np.float32(3.0), np.float32(10.0)
        elif ck == "diag":
            c1 = r2.uniform(0.5, 3, L).astype(np.float32)
            c2 = r2.uniform(1, 8, L).astype(np.float32)
        else:
            c1 = (-2 * np.eye(L) + r2.normal(0, 0.3, (L, L))).astype(np.float32)
            c2 = (-5 * np.eye(L) + r2.normal(0, 0.5, (L, L))).astype(np.float32)
        crf = po.RefCRF(W=W, H=H, L=L, identity_unary=True)
        crf.set_unary(U)
        crf.add_gaussian(3.0, 2.0, float(c1) if ck == "potts" else c1, ktype=kt1, ntype=nt)
        crf.add_bilateral(20.0, 30.0, 13.0, 17.0, 11.0, img, float(c2) if ck == "potts" else c2, ktype=kt2,
                          ntype=nt)
        kp = crf.kernel_params()
        if mode == "scaled":
            segs, j = [], 0
            for kt, d in ((kt1, 2), (kt2, 5)):
                if kt == 1:
                    segs.append(r2.uniform(0.7, 1.4, d))
                elif kt == 2:
                    segs.append((np.eye(d) + r2.normal(0, 0.1, (d, d))).T.ravel())
            kp = np.concatenate(segs).astype(np.float32)
            crf.set_kernel_params(kp)
        r, ug, cg, kg = crf.gradient(n, gt, kernel=True)
        p = "c%d_" % i
        out.update({p + "img": img, p + "U": U, p + "gt": gt, p + "c1": np.atleast_1d(c1), p + "c2": np.atleast_1d(c2),
                    p + "kparams": kp, p + "obj": np.float64(r), p + "unary_grad": ug, p + "compat_grad": cg,
                    p + "kernel_grad": kg, p + "Q": crf.inference(n),
                    p + "meta": np.array([W, H, L, ["potts", "diag", "matrix"].index(ck), kt1, kt2, nt, n],
                                         np.int32)})
    out["ncases"] = np.int32(len(KG_CASES))
    np.savez_compressed(os.path.join(HERE, "kgrad_small.npz"), **out)


OBJ_CASES = [
    # (objective, param, class weights or None, W, H, L, n_iter)
    (1, 0.0, None, 15, 11, 3, 3),
    (1, 1.0, None, 15, 11, 3, 3),
    (1, 0.0, [1.0, 2.0, 0.5, 3.0, 1.5], 13, 9, 5, 2),
    (2, 0.0, None, 15, 11, 3, 3),
    (2, 0.0, None, 16, 12, 21, 2),
    (0, 0.05, None, 15, 11, 3, 3),
]


def learning_small():
    """DenseCRF::gradient with the Hamming and IntersectionOverUnion objectives, and
    with a LogisticUnaryEnergy (its own unary gradient sum_t b f^T)."""
    out = {}
    for i, (obj, param, cw, W, H, L, n) in enumerate(OBJ_CASES):
        img, U, gt = syn.make_image(W, H, L=L, sites=4, seed=900 + i)
        crf = po.RefCRF(W=W, H=H, L=L, identity_unary=True)
        crf.set_unary(U)
        crf.add_gaussian(3.0, 2.0, 3.0, ntype=3)
        crf.add_bilateral(20.0, 30.0, 13.0, 17.0, 11.0, img, 10.0, ntype=3)
        cwa = None if cw is None else np.asarray(cw, np.float32)
        r, ug, cg, kg = crf.gradient_obj(n, gt, obj, param, cwa, kernel=True)
        p = "o%d_" % i
        out.update({p + "img": img, p + "U": U, p + "gt": gt, p + "obj": np.float64(r), p + "unary_grad": ug,
                    p + "compat_grad": cg, p + "kernel_grad": kg,
                    p + "class_weight": np.zeros(0, np.float32) if cwa is None else cwa,
                    p + "meta": np.array([obj, W, H, L, n], np.int32), p + "param": np.float32(param)})
    out["nobj"] = np.int32(len(OBJ_CASES))
    for i, (W, H, L, F, n) in enumerate([(14, 10, 3, 5, 3), (17, 9, 21, 13, 2)]):
        rng = np.random.default_rng(950 + i)
        img, _, gt = syn.make_image(W, H, L=L, sites=4, seed=950 + i)
        Lm = (rng.standard_normal((L, F)) * 0.5).astype(np.float32)
        f = rng.standard_normal((F, W * H)).astype(np.float32)
        crf = po.RefCRF(W=W, H=H, L=L)
        crf.set_logistic_unary(Lm, f)
        crf.add_gaussian(3.0, 2.0, 3.0, ntype=3)
        crf.add_bilateral(20.0, 30.0, 13.0, 17.0, 11.0, img, 10.0, ntype=3)
        r, ugp, cg, _ = crf.gradient_obj(n, gt, 0, 0.0, None, n_unary=L * F)
        p = "l%d_" % i
        out.update({p + "img": img, p + "gt": gt, p + "Lm": Lm, p + "f": f, p + "obj": np.float64(r),
                    p + "unary_param_grad": ugp, p + "compat_grad": cg, p + "params": crf.unary_params(),
                    p + "Q": crf.inference(n), p + "meta": np.array([W, H, L, F, n], np.int32)})
    out["nlog"] = np.int32(2)
    # DenseCRF::unaryEnergy / pairwiseEnergy of a labelling (labels outside [0, L) give 0)
    W, H, L = 19, 13, 5
    img, U, gt = syn.make_image(W, H, L=L, sites=5, seed=977)
    crf = po.RefCRF(W=W, H=H, L=L)
    crf.set_unary(U)
    crf.add_gaussian(3.0, 2.0, 3.0, ntype=3)
    crf.add_bilateral(20.0, 30.0, 13.0, 17.0, 11.0, img, np.arange(1, L + 1, dtype=np.float32), ntype=1)
    lab = np.random.default_rng(978).integers(-1, L + 1, W * H).astype(np.int16)
    out.update({"e_img": img, "e_U": U, "e_labels": lab, "e_unary": crf.unary_energy(lab),
                "e_pair0": crf.pairwise_energy(lab, 0), "e_pair1": crf.pairwise_energy(lab, 1),
                "e_pair": crf.pairwise_energy(lab, -1), "e_meta": np.array([W, H, L], np.int32)})
    np.savez_compressed(os.path.join(HERE, "learning_small.npz"), **out)


def simple_10x10():
    # the reference suite's fixture, pydensecrf/tests/test_dcrf.py:8-26
    unary1 = np.zeros((10, 10), dtype=np.float32)
    unary1[:, [0, -1]] = unary1[[0, -1], :] = 1
    unary2 = np.zeros((10, 10), dtype=np.float32)
    unary2[4:7, 4:7] = 1
    unary = np.vstack([unary1.flat, unary2.flat])
    unary = (unary + 1) / (np.sum(unary, axis=0) + 2)
    img = np.zeros((10, 10, 3), dtype=np.uint8)
    img[2:8, 2:8, :] = 255
    U = -np.log(unary).astype(np.float32)
    d2 = po.RefCRF(W=10, H=10, L=2)
    d2.set_unary(U)
    d2.add_bilateral(2.0, 2.0, 2.0, 2.0, 2.0, img, 3.0)
    Q2d = d2.inference(10)
    d = po.RefCRF(N=100, L=2)
    d.set_unary(U)
    feats = utils.create_pairwise_bilateral(sdims=(2, 2), schan=2, img=img, chdim=2)
    d.add_pairwise(feats, 3.0)
    Qg = d.inference(10)
    np.savez_compressed(os.path.join(HERE, "simple_10x10.npz"), U=U, img=img, feats=feats, Q2d=Q2d, Qgeneric=Qg)


def image_workload(name, cfg, fname):
    img, U, gt = syn.make_image(**cfg)
    W, H, L, n = cfg["W"], cfg["H"], cfg["L"], cfg["iters"]
    crf = po.RefCRF(W=W, H=H, L=L)
    crf.set_unary(U)
    crf.add_gaussian(3.0, 3.0, 3.0)
    crf.add_bilateral(80.0, 80.0, 13.0, 13.0, 13.0, img, 10.0)
    Q = crf.inference(n)
    rng = np.random.default_rng(7)
    idx = np.sort(rng.choice(W * H, 4096, replace=False))
    # lattice fingerprints from the standalone lattice probe on the same features
    xs, ys = np.meshgrid(np.arange(W), np.arange(H))
    fg = np.stack([xs.ravel().astype(np.float32) / np.float32(3.0), ys.ravel().astype(np.float32) / np.float32(3.0)])
    latg = po.RefLattice(fg).arrays
    np.savez_compressed(os.path.join(HERE, fname), W=W, H=H, L=L, iters=n, idx=idx, Q_sample=Q[:, idx],
                        map_sha=sha(Q.argmax(0).astype(np.int16)), unary_sha=sha(U), image_sha=sha(img),
                        V_gaussian=np.int32(latg.V), offset_gaussian_sha=sha(latg.offset[:W * H]))


def batch_64():
    """The 64-image batch (synthetic.image_batch, 10 iterations, Gaussian sxy=3 w=3 +
    bilateral sxy=80 srgb=13 w=10, Potts): the reference's Q on a 4096-pixel sample and
    its MAP hash for three instances, the smallest, the median and the largest."""
    sizes, imgs, Us = syn.image_batch(64)
    Ns = np.array([w * h for w, h in sizes])
    order = np.argsort(Ns, kind="stable")
    picks = [int(order[0]), int(order[len(order) // 2]), int(order[-1])]
    out = {"sizes": np.array(sizes, np.int32), "picks": np.array(picks, np.int32), "iters": np.int32(10)}
    for j, i in enumerate(picks):
        w, h = sizes[i]
        crf = po.RefCRF(W=w, H=h, L=21)
        crf.set_unary(Us[i])
        crf.add_gaussian(3.0, 3.0, 3.0)
        crf.add_bilateral(80.0, 80.0, 13.0, 13.0, 13.0, imgs[i], 10.0)
        Q = crf.inference(10)
        idx = np.sort(np.random.default_rng(11 + j).choice(w * h, 4096, replace=False))
        p = "p%d_" % j
        out.update({p + "idx": idx, p + "Q_sample": Q[:, idx], p + "map_sha": sha(Q.argmax(0).astype(np.int16)),
                    p + "unary_sha": sha(Us[i]), p + "image_sha": sha(imgs[i])})
    np.savez_compressed(os.path.join(HERE, "batch_64.npz"), **out)


def volumes_8():
    """The 8-volume batch (synthetic.volume_batch, L = 4, DiagonalCompatibility [1,2,2,3]
    per term, 10 iterations): for the first and the largest volume, the reference's Q on
    a 4096-voxel sample and its MAP hash, and DenseCRF::gradient (LogLikelihood of the
    true labels, identity-gradient unary) on that volume alone: objective, unary gradient
    on the sample, its largest magnitude over all entries, compatibility gradient; and
    FP64 central differences of the cross-entropy over the compatibility entries."""
    sizes, fgs, fbs, Us, gts = syn.volume_batch(8)
    picks = [0, int(np.argmax(sizes))]
    out = {"sizes": np.array(sizes, np.int64), "picks": np.array(picks, np.int32), "iters": np.int32(10)}
    for j, i in enumerate(picks):
        crf = po.RefCRF(N=sizes[i], L=4, identity_unary=True)
        crf.set_unary(Us[i])
        crf.add_pairwise(fgs[i], syn.VOLUME_COMPAT)
        crf.add_pairwise(fbs[i], syn.VOLUME_COMPAT)
        Q = crf.inference(10)
        r, ug, cg = crf.gradient(10, gts[i])
        # FP64 central differences (h = 1e-3) of the cross-entropy over the 8 Diagonal
        # compatibility entries, from the oracle's FP64 twin on the same lattices
        o = po.OracleCRF(N=sizes[i], L=4)
        o.set_unary(Us[i])
        o.add_pairwise(fgs[i], syn.VOLUME_COMPAT)
        o.add_pairwise(fbs[i], syn.VOLUME_COMPAT)
        raw = np.concatenate([syn.VOLUME_COMPAT, syn.VOLUME_COMPAT]).astype(np.float64)
        U64, fd = Us[i].astype(np.float64), []
        for k in range(raw.size):
            a, b = raw.copy(), raw.copy()
            a[k] += 1e-3
            b[k] -= 1e-3
            fd.append((o.ce_fp64(10, gts[i], U64, a, None) - o.ce_fp64(10, gts[i], U64, b, None)) / 2e-3)
        idx = np.sort(np.random.default_rng(21 + j).choice(sizes[i], 4096, replace=False))
        p = "p%d_" % j
        out.update({p + "idx": idx, p + "Q_sample": Q[:, idx], p + "map_sha": sha(Q.argmax(0).astype(np.int16)),
                    p + "unary_sha": sha(Us[i]), p + "obj": np.float64(r), p + "unary_grad_sample": ug[:, idx],
                    p + "unary_grad_absmax": np.float32(np.abs(ug).max()), p + "compat_grad": cg,
                    p + "compat_fd": np.array(fd)})
    np.savez_compressed(os.path.join(HERE, "volumes_8.npz"), **out)


# FP64 central differences of the 512x512 cross-entropy (workers of a process pool)
_FD = {}
FD_ITERS, FD_H = 5, 1e-3


def _fd_init():
    img, U, gt, c1, c2, kw = syn.gradient_problem_512()
    o = po.OracleCRF(W=512, H=512, L=21)
    o.set_unary(U)
    o.add_gaussian(3.0, 3.0, c1)
    o.add_bilateral(80.0, 80.0, 13.0, 13.0, 13.0, img, c2)
    _FD.update(o=o, U=U.astype(np.float64), gt=gt, raw=np.concatenate([c1.ravel(), c2.ravel()]).astype(np.float64),
               kw=kw.astype(np.float64), L=21)


def _fd_ce(kind, idx, delta):
    c = _FD
    U, raw, kw = c["U"], c["raw"], c["kw"]
    if kind == "u":
        U = U.copy()
        U.flat[idx] += delta
    elif kind == "c":  # one unordered pair (a, b) of term t: m_ab and m_ba move together
        t, a, b = idx
        L = c["L"]
        raw = raw.copy()
        raw[t * L * L + a * L + b] += delta
        if a != b:
            raw[t * L * L + b * L + a] += delta
    elif kind == "k":
        kw = kw.copy()
        kw[idx] += delta
    return c["o"].ce_fp64(FD_ITERS, c["gt"], U, raw, kw)


def _fd_job(job):
    kind, idx = job
    if kind == "0":
        return _fd_ce("0", None, 0.0)
    return (_fd_ce(kind, idx, FD_H) - _fd_ce(kind, idx, -FD_H)) / (2 * FD_H)


def grad_512x512(jobs=None):
    """The differentiable 512x512x21 workload (synthetic.gradient_problem_512, 5
    iterations).

    Reference: DenseCRF::gradient (LogLikelihood of the true labels, identity-gradient
    unary) with MatrixCompatibility w_k m_k (the term weight folded into the matrix):
    objective, dL/dU on 64 entries and its largest magnitude over all entries, the
    packed compatibility gradient of both terms.  The 64 unary entries are drawn (seed 7)
    32 among those whose reference gradient is at least 1% of the largest and 32 among
    those between 1e-4 and 1% of it: most pixels are saturated, and their entries (~1e-11
    of a largest ~6e-6) would only test rounding.

    FP64 central differences (h = 1e-3) of the mean cross-entropy with the oracle on the
    same lattices, over the parameters of the autograd layer (raw m_k and w_k): the 64
    unary entries, every unordered compatibility pair (a, b) of each term (m_ab and m_ba
    moved together, so the quotient is dCE/dm_ab + dCE/dm_ba off the diagonal), and the
    two term weights; plus the FP64 cross-entropy itself."""
    import multiprocessing as mp
    img, U, gt, c1, c2, kw = syn.gradient_problem_512()
    L = 21
    crf = po.RefCRF(W=512, H=512, L=L, identity_unary=True)
    crf.set_unary(U)
    crf.add_gaussian(3.0, 3.0, (kw[0] * c1).astype(np.float32))
    crf.add_bilateral(80.0, 80.0, 13.0, 13.0, 13.0, img, (kw[1] * c2).astype(np.float32))
    r, ug, cg = crf.gradient(FD_ITERS, gt)
    mag = np.abs(ug.ravel()) / np.abs(ug).max()
    rng = np.random.default_rng(7)
    sel = np.sort(np.concatenate([rng.choice(np.flatnonzero(mag >= 1e-2), 32, replace=False),
                                  rng.choice(np.flatnonzero((mag >= 1e-4) & (mag < 1e-2)), 32, replace=False)]))
    pairs = [(t, a, b) for t in range(2) for a in range(L) for b in range(a, L)]
    work = [("0", None)] + [("u", int(i)) for i in sel] + [("c", p) for p in pairs] + [("k", 0), ("k", 1)]
    with mp.Pool(jobs or os.cpu_count() or 1, initializer=_fd_init) as pool:
        vals = np.array(pool.map(_fd_job, work, chunksize=4))
    nu = len(sel)
    np.savez_compressed(os.path.join(HERE, "grad_512x512.npz"), sel=sel, ref_obj=np.float64(r),
                        ref_unary_grad_sel=ug.ravel()[sel], ref_unary_grad_absmax=np.float32(np.abs(ug).max()),
                        ref_compat_grad=cg, ce_fp64=np.float64(vals[0]), fd_unary=vals[1:1 + nu],
                        fd_pairs=vals[1 + nu:1 + nu + len(pairs)], pairs=np.array(pairs, np.int32),
                        fd_kweights=vals[1 + nu + len(pairs):], unary_sha=sha(U), image_sha=sha(img),
                        iters=np.int32(FD_ITERS), h=np.float64(FD_H))


def lbfgs_small():
    """minimizeLBFGS (optimization.cpp:68-103 over liblbfgs 1.10) from the reference
    build: every point it evaluates and its result, on the extended Rosenbrock function
    from two starts, and on CRF learning energies (tests/helpers.py crf_learning_energy:
    -LogLikelihood + l2 / 2 ||x||^2 evaluated with the reference's DenseCRF::gradient)
    with one restart: "crf" over Diagonal compatibilities and DIAG kernel parameters,
    "crfc" over the compatibilities alone (an energy smooth in its parameters)."""
    from tests.helpers import crf_learning_energy, crf_learning_model, rosenbrock
    out = {}

    def traced(fun, trace):
        def f(x):
            trace.append(x.copy())
            return fun(x)
        return f

    for i, x0 in enumerate([np.array([-1.2, 1.0] * 3, np.float32),
                            np.array([3.0, -2.0, 0.5, 4.0, -1.0, 1.0, 2.0, 2.0], np.float32)]):
        tr = []
        out["r%d_x" % i] = po.ref_minimize_lbfgs(x0, traced(rosenbrock, tr), 0)
        out["r%d_x0" % i], out["r%d_trace" % i] = x0, np.array(tr)
    for name, kernel in (("crf", True), ("crfc", False)):
        crf, gt = crf_learning_model(po.RefCRF, identity_unary=True)
        x0 = np.concatenate([crf.compat_params()] + ([crf.kernel_params()] if kernel else [])).astype(np.float32)
        tr, vals = [], []
        fun = crf_learning_energy(crf, gt, kernel=kernel)

        def rec(x, fun=fun, tr=tr, vals=vals):
            tr.append(x.copy())
            f, g = fun(x)
            vals.append(f)
            return f, g

        out[name + "_x"] = po.ref_minimize_lbfgs(x0, rec, 1)
        out[name + "_x0"], out[name + "_trace"], out[name + "_values"] = x0, np.array(tr), np.array(vals)
    np.savez_compressed(os.path.join(HERE, "lbfgs_small.npz"), **out)


if __name__ == "__main__":
    if not po.ref_available():
        sys.exit("build the reference first: make -C oracle ref")
    writers = {"lattice_small": lattice_small, "crf_small": crf_small, "kgrad_small": kgrad_small,
               "learning_small": learning_small,
               "simple_10x10": simple_10x10,
               "image_320x213": lambda: image_workload("320x213", syn.IMAGE_320x213, "image_320x213.npz"),
               "image_1024x768": lambda: image_workload("1024x768", syn.IMAGE_1024x768, "image_1024x768.npz"),
               "batch_64": batch_64, "volumes_8": volumes_8, "grad_512x512": grad_512x512,
               "lbfgs_small": lbfgs_small}
    # python tests/golden/make_golden.py [fixture ...]: all of them by default
    for name in sys.argv[1:] or list(writers):
        writers[name]()
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))
