"""FP64 central-difference check of the adjoint at 512x512x21 (not collected by pytest).

Workload: DenseCRF2D 512x512, 21 labels, 5 iterations, MatrixCompatibility per term
(-I + N(0, 0.1), learnable), learnable term weights (3, 10), Gaussian sxy=3 +
bilateral sxy=80 srgb=13, cross-entropy of the final Q against the generator's true
labels.

    python tests/fd_check_512x512.py --gpu OUT.npz     # on a GPU: analytic gradients
                                                       # through pydensecrf_amd.autograd
    python tests/fd_check_512x512.py --fd OUT.npz [-j 8]  # on the CPU: FP64 central
                                                       # differences (h = 1e-3) with the
                                                       # oracle on the same lattices
                                                       # (saved as OUT_fd.npz; --fd-values
                                                       # OUT_fd.npz reuses them)
The --fd step checks 64 seeded unary entries, every distinct compatibility entry
(m_ab and m_ba share one value of W = (m + m^T)/2, so each unordered pair is
differenced once and compared with both gradient entries) and the two term weights,
and prints the largest error relative to max(|fd|, 1e-4 max|fd|) (bar 1e-3).
"""
import argparse
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

N_ITER, H_STEP = 5, 1e-3


def problem():
    from pydensecrf_amd import synthetic as syn
    cfg = syn.IMAGE_512x512
    img, U, gt = syn.make_image(**cfg)
    L = cfg["L"]
    rng = np.random.default_rng(1003)
    c1 = (-np.eye(L) + rng.normal(0, 0.1, (L, L))).astype(np.float32)
    c2 = (-np.eye(L) + rng.normal(0, 0.1, (L, L))).astype(np.float32)
    kw = np.array([3.0, 10.0], np.float32)
    sel = np.random.default_rng(7).choice(L * cfg["W"] * cfg["H"], 64, replace=False)
    return cfg, img, U, gt, c1, c2, kw, sel


def run_gpu(out):
    import torch

    from pydensecrf_amd import capi
    from pydensecrf_amd.autograd import CRFLayer
    cfg, img, U, gt, c1, c2, kw, sel = problem()
    W, H, L = cfg["W"], cfg["H"], cfg["L"]
    m = capi.Model.create_2d(W, H, L)
    m.add_gaussian(3.0, 3.0, c1)
    m.add_bilateral(80.0, 80.0, 13.0, 13.0, 13.0, img, c2)
    layer = CRFLayer(m, N_ITER, [c1, c2], kweights=kw)
    Ut = torch.tensor(np.ascontiguousarray(U.T), device="cuda", requires_grad=True)
    Q = layer(Ut)
    gtt = torch.tensor(gt.astype(np.int64), device="cuda")
    loss = torch.nn.functional.nll_loss(torch.log(Q), gtt)
    loss.backward()
    dU = Ut.grad.cpu().numpy().T.ravel()  # (L, N) order
    np.savez(out, loss=float(loss), dU_sel=dU[sel], sel=sel, dcompat=layer.compat.grad.cpu().numpy(),
             dkw=layer.kweights.grad.cpu().numpy())
    print("gpu: loss %.8f, gradients written to %s" % (float(loss), out), flush=True)


_CTX = {}


def _init():
    from oracle import pyoracle as po
    cfg, img, U, gt, c1, c2, kw, sel = problem()
    W, H, L = cfg["W"], cfg["H"], cfg["L"]
    o = po.OracleCRF(W=W, H=H, L=L)
    o.set_unary(U)
    o.add_gaussian(3.0, 3.0, c1)
    o.add_bilateral(80.0, 80.0, 13.0, 13.0, 13.0, img, c2)
    raw = np.concatenate([c1.ravel(), c2.ravel()]).astype(np.float64)
    _CTX.update(o=o, U=U.astype(np.float64), gt=gt, raw=raw, kw=kw.astype(np.float64))


def _ce(kind, idx, delta):
    c = _CTX
    U, raw, kw = c["U"], c["raw"], c["kw"]
    if kind == "u":
        U = U.copy()
        U.flat[idx] += delta
    elif kind == "c":  # one unordered pair (a, b) of term t: both m_ab and m_ba move
        t, a, b = idx
        L = c["U"].shape[0]
        raw = raw.copy()
        raw[t * L * L + a * L + b] += delta
        if a != b:
            raw[t * L * L + b * L + a] += delta
    else:
        kw = kw.copy()
        kw[idx] += delta
    return c["o"].ce_fp64(N_ITER, c["gt"], U, raw, kw)


def _fd(job):
    kind, idx = job
    return (_ce(kind, idx, H_STEP) - _ce(kind, idx, -H_STEP)) / (2 * H_STEP)


def run_fd(path, jobs, fd_values=None):
    import multiprocessing as mp
    g = np.load(path)
    cfg = problem()[0]
    L = cfg["L"]
    work = [("u", int(i)) for i in g["sel"]]
    pairs = [(t, a, b) for t in range(2) for a in range(L) for b in range(a, L)]
    work += [("c", p) for p in pairs] + [("k", 0), ("k", 1)]
    if fd_values:  # difference quotients of an earlier run (they depend on the problem only)
        fd = np.load(fd_values)["fd"]
        assert fd.shape == (len(work),)
    else:
        with mp.Pool(jobs, initializer=_init) as pool:
            fd = pool.map(_fd, work, chunksize=4)
        fd = np.array(fd)
    nu = len(g["sel"])
    # analytic values in the same order; a pair (a, b) moves m_ab and m_ba together, so
    # its difference quotient is dL/dm_ab + dL/dm_ba (= 2 dL/dm_ab off the diagonal)
    dc = g["dcompat"]
    an = list(g["dU_sel"])
    for t, a, b in pairs:
        v = dc[t * L * L + a * L + b]
        w = dc[t * L * L + b * L + a]
        assert abs(v - w) <= 1e-6 * max(abs(v), 1e-12) + 1e-12, "W symmetry broken in the gradient"
        an.append(v + w if a != b else v)
    an += list(g["dkw"])
    an = np.array(an)
    np.savez(os.path.splitext(path)[0] + "_fd.npz", fd=fd, analytic=an)
    # mixed tolerance: relative to the entry, floored at 1e-4 of the largest gradient
    # (fp32 accumulation leaves ~1e-8 of the gradient scale on entries near zero)
    rel = np.abs(an - fd) / np.maximum(np.abs(fd), 1e-4 * np.abs(fd).max())
    parts = {"unary": rel[:nu], "compat": rel[nu:nu + len(pairs)], "kweights": rel[nu + len(pairs):]}
    for k, r in parts.items():
        print("%-8s n=%4d  max rel err %.3g  median %.3g" % (k, r.size, r.max(), np.median(r)))
    labels = ["u%d" % i for i in g["sel"]] + ["c%d:%d,%d" % p for p in pairs] + ["kw0", "kw1"]
    for i in np.argsort(-rel)[:8]:
        print("  worst %-12s analytic % .6e  fd % .6e  rel %.3g  |err|/max|fd| %.3g"
              % (labels[i], an[i], fd[i], rel[i], abs(an[i] - fd[i]) / np.abs(fd).max()))
    print("overall max rel err %.3g (bar 1e-3)" % rel.max())
    return rel.max()


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpu", metavar="OUT")
    ap.add_argument("--fd", metavar="GRADS")
    ap.add_argument("-j", type=int, default=os.cpu_count() or 1)
    ap.add_argument("--fd-values", metavar="NPZ", help="reuse the difference quotients saved by an earlier --fd run")
    a = ap.parse_args()
    if a.gpu:
        run_gpu(a.gpu)
    if a.fd:
        sys.exit(0 if run_fd(a.fd, a.j, a.fd_values) <= 1e-3 else 1)
