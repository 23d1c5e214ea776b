#!/usr/bin/env python3
"""Dense-CRF mean-field inference benchmark on one MI355X.

Workload: DenseCRF2D 1024x768, 21 labels, 10 mean-field iterations, Gaussian
(sxy=3, Potts w=3) + bilateral (sxy=80, srgb=13, Potts w=10) pairwise terms,
DIAG_KERNEL, NORMALIZE_SYMMETRIC, on a seeded synthetic Voronoi image and unary
(pydensecrf_amd.synthetic).  A step is one inference call (10 iterations) over the
image with the unary and the lattices resident in HBM; Q is written to a device
buffer.  value = Mpixel*iterations/s over the timed steps.

Also reported on the same line: the blur-pass roofline (algorithmic bytes of the
timed blur launches / their HIP-event time, against 8 TB/s), lattice-build time, and
the reference CPU path (oracle/_ref, the reference's own C++ built from its sources;
the oracle restatement if that build is absent) timed on the host on the same
workload.

    python bench.py [--steps K] [--warmup W] [--iters I] [--no-cpu-baseline]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
WORKLOAD = "DenseCRF2D 1024x768, 21 labels, 10 mean-field iterations, Gaussian(sxy=3,w=3) + Bilateral(sxy=80,srgb=13,w=10), Potts, NORMALIZE_SYMMETRIC"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(img, U, W, H, L, iters):
    """Reference CPU path on the same workload (single thread, -O2, no -march)."""
    import numpy as np  # noqa: F401
    from oracle import pyoracle as po
    if po.ref_available():
        crf, kind = po.RefCRF(W=W, H=H, L=L), "reference"
    else:
        crf, kind = po.OracleCRF(W=W, H=H, L=L), "port"
    crf.set_unary(U)
    t0 = time.perf_counter()
    crf.add_gaussian(3.0, 3.0, 3.0)
    crf.add_bilateral(80.0, 80.0, 13.0, 13.0, 13.0, img, 10.0)
    t1 = time.perf_counter()
    crf.inference(iters)
    t2 = time.perf_counter()
    N = W * H
    return {
        "value": N * iters / (t2 - t1) / 1e6,
        "unit": "Mpixel*iters/s",
        "cores": 1,
        "kind": kind,
        "sample": "full %dx%dx%d image, %d iterations, one run; lattice build %.2f s excluded" % (W, H, L, iters,
                                                                                                  t1 - t0),
        "inference_s": t2 - t1,
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--iters", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the mean-field path runs on a single MI355X here "
              "(multi-GPU sharding is described in DESIGN.md, not built); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from pydensecrf_amd import capi
    from pydensecrf_amd import synthetic as syn

    cfg = syn.IMAGE_1024x768
    W, H, L = cfg["W"], cfg["H"], cfg["L"]
    iters = args.iters
    N = W * H
    img, U, _ = syn.make_image(**cfg)
    torch.cuda.init()
    dev = torch.device("cuda:0")

    m = capi.Model.create_2d(W, H, L)
    m.set_unary(U)
    t0 = time.perf_counter()
    m.add_gaussian(3.0, 3.0, 3.0)
    t1 = time.perf_counter()
    m.add_bilateral(80.0, 80.0, 13.0, 13.0, 13.0, img, 10.0)
    t2 = time.perf_counter()
    Vg, Vb = m.term_info(0)[1], m.term_info(1)[1]
    log("lattices: gaussian V=%d (%.1f ms), bilateral V=%d (%.1f ms)" % (Vg, (t1 - t0) * 1e3, Vb, (t2 - t1) * 1e3))

    Q = torch.empty(N * L, device=dev, dtype=torch.float32)
    stream = torch.cuda.current_stream(dev)

    def step():
        m.inference_device(iters, Q.data_ptr(), None, stream.cuda_stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    m.enable_timing(True)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(args.steps):
        step()
    e1.record(stream)
    torch.cuda.synchronize()
    ms_total = e0.elapsed_time(e1)
    launches, blur_ms, blur_bytes = m.blur_stats()
    m.enable_timing(False)
    ms_per_step = ms_total / args.steps
    value = N * iters * args.steps / (ms_total / 1e3) / 1e6

    # sanity: Q rows are distributions
    q = Q.view(N, L)
    row_err = float((q.sum(1) - 1).abs().max().item())
    if not row_err < 1e-4:
        raise RuntimeError("Q rows do not sum to one (max error %g)" % row_err)

    blur_gbs = blur_bytes / (blur_ms / 1e3) / 1e9 if blur_ms > 0 else 0.0
    iter_bytes = m.iteration_bytes()
    iter_gbs = iter_bytes * iters / (ms_per_step / 1e3) / 1e9
    out = {
        "metric": "Mpixel*iters/s, dense-CRF mean-field forward at 1024x768x21",
        "value": round(value, 3),
        "unit": "Mpixel*iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded Voronoi image and unary, pydensecrf_amd.synthetic)",
        "config": {"workload": WORKLOAD, "width": W, "height": H, "labels": L, "iterations": iters,
                   "vertices": {"gaussian": Vg, "bilateral": Vb}},
        "roofline": {
            "bound": "hbm",
            "kernel": "k_blur (blur pass)",
            "achieved": round(blur_gbs, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(blur_gbs / HBM_PEAK_GBS, 4),
            "traffic": None,
            "launches": launches,
            "avg_launch_us": round(blur_ms * 1e3 / max(launches, 1), 3),
            "bytes_per_launch": round(blur_bytes / max(launches, 1)),
        },
        "iteration": {"ms": round(ms_per_step / iters, 4), "algorithmic_bytes": round(iter_bytes),
                      "achieved_GBs": round(iter_gbs, 2), "frac": round(iter_gbs / HBM_PEAK_GBS, 4)},
        "lattice_build_ms": {"gaussian": round((t1 - t0) * 1e3, 2), "bilateral": round((t2 - t1) * 1e3, 2)},
    }
    if not args.no_cpu_baseline:
        log("timing the reference CPU path on the same workload ...")
        cb = cpu_baseline(img, U, W, H, L, iters)
        out["cpu_baseline"] = {k: cb[k] for k in ("value", "unit", "cores", "kind", "sample")}
        out["cpu_baseline"]["value"] = round(cb["value"], 4)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
