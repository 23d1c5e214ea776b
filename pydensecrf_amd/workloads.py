"""Fixed workloads for profiling the paths beside the inference benchmark.

    python -m pydensecrf_amd.workloads kernel-gradient [--steps S]
        DenseCRF2D 512x512, 21 labels, 5 iterations, Matrix compatibilities, learnable
        DIAG kernel parameters: S forward+backward passes through the torch.autograd
        layer (the kernel-parameter gradient in every backward), after two warm-up passes.
    python -m pydensecrf_amd.workloads kernel-learning [--steps S]
        The same model and layer with the kernel parameters changed after every step (a
        sign step of 0.1 %): each step rebuilds both lattices, the layout and the
        kernel gradient's filter plan, then runs the forward+backward.
    python -m pydensecrf_amd.workloads logistic-unary [--reps R] [--features F]
        1024x768 pixels, 21 labels, a logistic unary with F features (default 64):
        R evaluations of U = L f on the matrix cores, with the parameters changed
        between evaluations so that each one recomputes U.

    python -m pydensecrf_amd.workloads fwd-bwd [--steps S] [--size 512|1024]
        The differentiable workloads through the torch.autograd layer: 512 = 512x512, 21
        labels, 5 iterations, Matrix compatibilities and term weights learnable; 1024 =
        1024x768, 21 labels, 10 iterations, Potts compatibilities and term weights
        learnable.  S forward+backward passes (cross-entropy of the true labels).
    python -m pydensecrf_amd.workloads kernel-rebuild [--reps R]
        DenseCRF2D 512x512, 21 labels, Gaussian (d = 2) + bilateral (d = 5) with DIAG
        kernel parameters: R changes of the parameter vector, each rebuilding both
        lattices (dcrf_set_kernel_params) and the model's layout.

    python -m pydensecrf_amd.workloads dropin [--reps R]
        The drop-in Cython API with host arrays, DenseCRF2D 1024x768, 21 labels,
        Gaussian (sxy=3, compat=3) + bilateral (sxy=80, srgb=13, compat=10): R images,
        each a new model (DenseCRF2D + setUnaryEnergy, both pairwise terms,
        inference(10)), after one warm-up image; then, on the last model, R calls each of
        inference(10), setUnaryEnergy + inference(10) and map(10).

Each prints one line with its host-clock time per repetition; run under rocprofv3 for
the per-kernel figures.
"""
import argparse
import time

import numpy as np


def kernel_gradient(steps):
    import torch

    from . import capi, synthetic as syn
    from .autograd import CRFLayer

    cfg = syn.IMAGE_512x512
    W, H, L = cfg["W"], cfg["H"], cfg["L"]
    img, U, gt = syn.make_image(**cfg)
    rng = np.random.default_rng(1003)
    c1 = (-np.eye(L) + rng.normal(0, 0.1, (L, L))).astype(np.float32)
    c2 = (-np.eye(L) + rng.normal(0, 0.1, (L, L))).astype(np.float32)
    m = capi.Model.create_2d(W, H, L)
    m.add_gaussian(3.0, 3.0, c1)
    m.add_bilateral(80.0, 80.0, 13.0, 13.0, 13.0, img, c2)
    layer = CRFLayer(m, 5, [c1, c2], kweights=[3.0, 10.0], learn_kernel=True)
    Ut = torch.tensor(np.ascontiguousarray(U.T), device="cuda", requires_grad=True)
    gtt = torch.tensor(gt.astype(np.int64), device="cuda")

    def step():
        Q = layer(Ut)
        loss = -torch.log(Q.gather(1, gtt[:, None])).mean()
        loss.backward()

    for _ in range(2):  # the second step is the first to accumulate into .grad (torch loads those kernels then)
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    print("kernel-gradient 512x512x21, 5 iterations: %.3f ms per forward+backward" % ms, flush=True)


def kernel_learning(steps):
    import torch

    from . import capi, synthetic as syn
    from .autograd import CRFLayer

    cfg = syn.IMAGE_512x512
    W, H, L = cfg["W"], cfg["H"], cfg["L"]
    img, U, gt = syn.make_image(**cfg)
    rng = np.random.default_rng(1003)
    c1 = (-np.eye(L) + rng.normal(0, 0.1, (L, L))).astype(np.float32)
    c2 = (-np.eye(L) + rng.normal(0, 0.1, (L, L))).astype(np.float32)
    m = capi.Model.create_2d(W, H, L)
    m.add_gaussian(3.0, 3.0, c1)
    m.add_bilateral(80.0, 80.0, 13.0, 13.0, 13.0, img, c2)
    layer = CRFLayer(m, 5, [c1, c2], kweights=[3.0, 10.0], learn_kernel=True)
    Ut = torch.tensor(np.ascontiguousarray(U.T), device="cuda", requires_grad=True)
    gtt = torch.tensor(gt.astype(np.int64), device="cuda")

    def step():
        Q = layer(Ut)  # new kernel parameters: both lattices and the layout are rebuilt
        loss = -torch.log(Q.gather(1, gtt[:, None])).mean()
        loss.backward()  # the kernel-gradient filter plan of the new lattices is built
        with torch.no_grad():
            layer.kparams.mul_(1.0 - 1e-3 * torch.sign(layer.kparams.grad))
            layer.kparams.grad = None

    for _ in range(2):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    print("kernel-learning 512x512x21, 5 iterations: %.3f ms per step (parameter change, forward+backward)" % ms,
          flush=True)


def fwd_bwd(steps, size):
    import torch

    from . import capi, synthetic as syn
    from .autograd import CRFLayer

    if size == 512:
        cfg = syn.IMAGE_512x512
        img, U, gt, c1, c2, kw = syn.gradient_problem_512()
        compat, n = [c1, c2], cfg["iters"]
    else:
        cfg = syn.IMAGE_1024x768
        img, U, gt = syn.make_image(**cfg)
        compat, kw, n = [3.0, 10.0], None, cfg["iters"]
    W, H, L = cfg["W"], cfg["H"], cfg["L"]
    m = capi.Model.create_2d(W, H, L)
    m.add_gaussian(3.0, 3.0, compat[0])
    m.add_bilateral(80.0, 80.0, 13.0, 13.0, 13.0, img, compat[1])
    layer = CRFLayer(m, n, compat, kweights=kw)
    Ut = torch.tensor(np.ascontiguousarray(U.T), device="cuda", requires_grad=True)
    gtt = torch.tensor(gt.astype(np.int64), device="cuda")

    def step():
        Q = layer(Ut)
        loss = -torch.log(Q.gather(1, gtt[:, None])).mean()
        loss.backward()

    for _ in range(2):  # the second step is the first to accumulate into .grad (torch loads those kernels then)
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    print("fwd-bwd %dx%dx%d, %d iterations: %.3f ms per forward+backward" % (W, H, L, n, ms), flush=True)


def logistic_unary(reps, F):
    from . import capi

    W, H, L = 1024, 768, 21
    N = W * H
    rng = np.random.default_rng(77)
    Lm = (rng.standard_normal((L, F)) * 0.2).astype(np.float32)
    f = rng.standard_normal((F, N)).astype(np.float32)
    m = capi.Model.create_2d(W, H, L)
    m.set_logistic_unary(Lm, f)
    Q = np.empty((N, L), np.float32)
    Qd = None
    try:
        import torch

        Qd = torch.empty((N, L), dtype=torch.float32, device="cuda")
    except Exception:
        pass
    p = m.unary_params()
    t = []
    for r in range(reps + 1):
        m.set_unary_params(p * np.float32(1.0 + 1e-3 * r))
        t0 = time.perf_counter()
        if Qd is not None:
            m.inference_device(0, Qd.data_ptr())  # U on the device, softmax(-U) into Qd
            torch.cuda.synchronize()
        else:
            Q[:] = m.inference(0).T
        t.append(time.perf_counter() - t0)
    ms = 1e3 * float(np.median(t[1:]))
    print("logistic-unary %dx%d, %d labels, %d features: %.3f ms per evaluation (U = L f and softmax)"
          % (W, H, L, F, ms), flush=True)


def kernel_rebuild(reps):
    from . import capi, synthetic as syn

    cfg = syn.IMAGE_512x512
    W, H, L = cfg["W"], cfg["H"], cfg["L"]
    img, U, _ = syn.make_image(**cfg)
    m = capi.Model.create_2d(W, H, L)
    m.set_unary(U)
    m.add_gaussian(3.0, 3.0, 3.0)
    m.add_bilateral(80.0, 80.0, 13.0, 13.0, 13.0, img, 10.0)
    import torch

    Qd = torch.empty((W * H, L), dtype=torch.float32, device="cuda")
    kp = m.kernel_params()

    split = [0.0, 0.0]

    def change(f):
        # new lattices, then the layout (built by the next call on the model: here a
        # 0-iteration inference, Q = softmax(-U))
        t0 = time.perf_counter()
        m.set_kernel_params(kp * np.float32(f))
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        m.inference_device(0, Qd.data_ptr())
        torch.cuda.synchronize()
        split[0] += t1 - t0
        split[1] += time.perf_counter() - t1

    change(1.001)
    split[:] = [0.0, 0.0]
    t0 = time.perf_counter()
    for r in range(reps):
        change(1.0 + 0.01 * (r + 1))
    ms = (time.perf_counter() - t0) * 1e3 / reps
    print("kernel-rebuild 512x512x21, d = 2 + 5: %.3f ms per parameter change (lattices %.3f ms, layout %.3f ms)"
          % (ms, split[0] * 1e3 / reps, split[1] * 1e3 / reps), flush=True)


def dropin(reps):
    from . import densecrf as dcrf, synthetic as syn

    W, H, L = 1024, 768, 21
    img, U, _ = syn.make_image(W, H, L=L, seed=0)
    U = np.ascontiguousarray(U, dtype=np.float32)
    img = np.ascontiguousarray(img)
    phases = {"create+unary": [], "gaussian": [], "bilateral": [], "inference": []}
    d = None
    for r in range(reps + 1):
        t0 = time.perf_counter()
        d = dcrf.DenseCRF2D(W, H, L)
        d.setUnaryEnergy(U)
        t1 = time.perf_counter()
        d.addPairwiseGaussian(sxy=3, compat=3)
        t2 = time.perf_counter()
        d.addPairwiseBilateral(sxy=80, srgb=13, rgbim=img, compat=10)
        t3 = time.perf_counter()
        Q = d.inference(10)
        t4 = time.perf_counter()
        if r:  # the first image warms the runtime up
            for k, v in zip(phases, (t1 - t0, t2 - t1, t3 - t2, t4 - t3)):
                phases[k].append(v)
    del Q
    per_image = sum(np.mean(v) for v in phases.values())
    print("dropin 1024x768x21, new model per image: %.2f ms (%s)" % (
        1e3 * per_image, ", ".join("%s %.2f" % (k, 1e3 * np.mean(v)) for k, v in phases.items())))
    reused = {"inference(10)": [], "setUnaryEnergy+inference(10)": [], "map(10)": []}
    for _ in range(reps):
        t0 = time.perf_counter()
        Q = d.inference(10)
        t1 = time.perf_counter()
        d.setUnaryEnergy(U)
        Q = d.inference(10)
        t2 = time.perf_counter()
        d.map(10)
        t3 = time.perf_counter()
        for k, v in zip(reused, (t1 - t0, t2 - t1, t3 - t2)):
            reused[k].append(v)
    print("dropin 1024x768x21, model reused: " + ", ".join("%s %.2f ms" % (k, 1e3 * np.mean(v))
                                                           for k, v in reused.items()))


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("workload", choices=["kernel-gradient", "kernel-learning", "logistic-unary", "kernel-rebuild",
                                         "fwd-bwd", "dropin"])
    ap.add_argument("--size", type=int, default=512, choices=[512, 1024])
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--features", type=int, default=64)
    a = ap.parse_args()
    if a.workload == "kernel-gradient":
        kernel_gradient(a.steps)
    elif a.workload == "kernel-learning":
        kernel_learning(a.steps)
    elif a.workload == "fwd-bwd":
        fwd_bwd(a.steps, a.size)
    elif a.workload == "kernel-rebuild":
        kernel_rebuild(a.reps)
    elif a.workload == "dropin":
        dropin(a.reps)
    else:
        logistic_unary(a.reps, a.features)


if __name__ == "__main__":
    main()
