"""Seeded synthetic workloads for dense-CRF mean-field benchmarks and parity tests.

Images are Voronoi partitions with a uniform random colour per region plus Gaussian
pixel noise; unaries are ``-log softmax`` of standard-normal logits with a +3 bump on
each region's label, where 30% of the regions carry a wrong label.  Volumes follow the
same recipe in 3-D with a scalar intensity.  Everything is numpy on the host.
"""
import numpy as np

# Named workloads: (width, height, labels, iterations, Voronoi sites, seed)
IMAGE_320x213 = dict(W=320, H=213, L=21, iters=5, sites=12, seed=1000)
IMAGE_1024x768 = dict(W=1024, H=768, L=21, iters=10, sites=48, seed=1001)
IMAGE_512x512 = dict(W=512, H=512, L=21, iters=5, sites=48, seed=1003)


def voronoi_labels(rng, shape, n_sites):
    """Nearest-site index per cell of a grid of ``shape`` (ties go to the lower index)."""
    coords = np.stack(np.meshgrid(*[np.arange(s, dtype=np.float64) for s in shape], indexing="ij"), -1)
    sites = np.stack([rng.uniform(0, s, n_sites) for s in shape], -1)
    flat = coords.reshape(-1, len(shape))
    best = np.zeros(flat.shape[0], dtype=np.int64)
    bestd = np.full(flat.shape[0], np.inf)
    for k in range(n_sites):
        dist = ((flat - sites[k]) ** 2).sum(-1)
        upd = dist < bestd
        best[upd] = k
        bestd[upd] = dist[upd]
    return best.reshape(shape)


def _region_labels(rng, n_sites, L, wrong_frac=0.3):
    true = rng.integers(0, L, n_sites)
    shown = true.copy()
    n_wrong = int(round(wrong_frac * n_sites)) if L > 1 else 0
    wrong = rng.choice(n_sites, n_wrong, replace=False) if n_wrong else np.array([], dtype=np.int64)
    for r in wrong:
        alt = rng.integers(0, L - 1)
        shown[r] = alt if alt < true[r] else alt + 1
    return true, shown


def _unary(rng, region, shown, L):
    n = region.size
    logits = rng.standard_normal((L, n))
    logits[shown[region.ravel()], np.arange(n)] += 3.0
    logits -= logits.max(0, keepdims=True)
    logp = logits - np.log(np.exp(logits).sum(0, keepdims=True))
    return np.ascontiguousarray((-logp).astype(np.float32))


def make_image(W, H, L=21, sites=12, seed=1000, **_):
    """Returns (image HxWx3 uint8, unary (L, W*H) float32, true labels (W*H,) int16)."""
    rng = np.random.default_rng(seed)
    region = voronoi_labels(rng, (H, W), sites)
    colours = rng.integers(0, 256, (sites, 3)).astype(np.float64)
    img = colours[region] + rng.normal(0.0, 8.0, (H, W, 3))
    img = np.clip(np.rint(img), 0, 255).astype(np.uint8)
    true, shown = _region_labels(rng, sites, L)
    U = _unary(rng, region, shown, L)
    gt = true[region.ravel()].astype(np.int16)
    return np.ascontiguousarray(img), U, gt


def make_volume(Z, Y, X, L=4, sites=20, seed=2000):
    """Returns (intensity ZxYxX float32, unary (L, N) float32, true labels (N,) int16)."""
    rng = np.random.default_rng(seed)
    region = voronoi_labels(rng, (Z, Y, X), sites)
    level = rng.uniform(0.0, 1.0, sites)
    vol = (level[region] + rng.normal(0.0, 0.05, region.shape)).astype(np.float32)
    true, shown = _region_labels(rng, sites, L)
    U = _unary(rng, region, shown, L)
    gt = true[region.ravel()].astype(np.int16)
    return vol, U, gt


def batch_image_sizes(n=64, lo=256, hi=1024, seed=1002):
    """(W, H) pairs drawn uniformly per side from [lo, hi]."""
    rng = np.random.default_rng(seed)
    return [(int(w), int(h)) for w, h in zip(rng.integers(lo, hi + 1, n), rng.integers(lo, hi + 1, n))]


def batch_volume_depths(n=8, lo=32, hi=64, seed=1004):
    rng = np.random.default_rng(seed)
    return [int(z) for z in rng.integers(lo, hi + 1, n)]


def image_batch(n=64, L=21):
    """The uneven image batch: sides U[256, 1024] (batch_image_sizes), 12 Voronoi sites,
    image i seeded 3000 + i.  Returns (sizes [(W, H)], images, unaries)."""
    sizes = batch_image_sizes(n)
    imgs, Us = [], []
    for i, (w, h) in enumerate(sizes):
        img, U, _ = make_image(w, h, L=L, sites=12, seed=3000 + i)
        imgs.append(img)
        Us.append(U)
    return sizes, imgs, Us


VOLUME_COMPAT = np.array([1, 2, 2, 3], np.float32)  # DiagonalCompatibility of both terms


def volume_features(vol):
    """Gaussian features (z, y, x) / 3 (d = 3) and bilateral features (z, y, x) / 40 with
    intensity / 0.1 (d = 4), each (d, Z*Y*X) float32 in (z, y, x) raster order."""
    Z, Y, X = vol.shape
    zz, yy, xx = np.meshgrid(np.arange(Z), np.arange(Y), np.arange(X), indexing="ij")
    pos = np.stack([zz, yy, xx]).reshape(3, -1).astype(np.float32)
    fg = pos / np.float32(3)
    fb = np.concatenate([pos / np.float32(40), vol.reshape(1, -1) / np.float32(0.1)])
    return np.ascontiguousarray(fg), np.ascontiguousarray(fb)


def volume_batch(n=8, L=4):
    """The volume batch: n volumes of 96x96xZ, Z ~ U{32..64} (batch_volume_depths), volume
    i seeded 2000 + i.  Returns (voxel counts, Gaussian features, bilateral features,
    unaries, true labels)."""
    sizes, fgs, fbs, Us, gts = [], [], [], [], []
    for i, Z in enumerate(batch_volume_depths(n)):
        vol, U, gt = make_volume(Z, 96, 96, L=L, seed=2000 + i)
        fg, fb = volume_features(vol)
        sizes.append(Z * 96 * 96)
        fgs.append(fg)
        fbs.append(fb)
        Us.append(U)
        gts.append(gt)
    return sizes, fgs, fbs, Us, gts


def gradient_problem_512():
    """The differentiable 512x512x21 workload (5 iterations): image, unary and true labels
    of IMAGE_512x512, MatrixCompatibility per term initialised to -I + N(0, 0.1) (seed
    1003; Gaussian term first), term weights (3, 10) for the Gaussian (sxy = 3) and the
    bilateral (sxy = 80, srgb = 13) term.  Returns (img, U, gt, c1, c2, kweights)."""
    cfg = IMAGE_512x512
    L = cfg["L"]
    img, U, gt = make_image(**cfg)
    rng = np.random.default_rng(1003)
    c1 = (-np.eye(L) + rng.normal(0, 0.1, (L, L))).astype(np.float32)
    c2 = (-np.eye(L) + rng.normal(0, 0.1, (L, L))).astype(np.float32)
    return img, U, gt, c1, c2, np.array([3.0, 10.0], np.float32)
