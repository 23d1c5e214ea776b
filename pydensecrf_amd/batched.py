"""Batched dense-CRF inference: many images (or volumes) of uneven size in one model.

All instances share the label count and the pairwise-term recipe; their lattices are
built in one pass (the instance id is part of every lattice key) and concatenated, so
every kernel of an iteration covers the whole batch in one launch.  Results per
instance are bit-identical to running each image through its own ``DenseCRF2D``.
``capture()`` records the whole n-iteration inference as a hipGraph in the library
(dcrf_capture_inference); ``replay()`` launches it.

    b = BatchedDenseCRF2D([(W0, H0), (W1, H1), ...], n_labels)
    b.setUnaryEnergy([U0, U1, ...])                      # each (L, W_i*H_i) float32
    b.addPairwiseGaussian(sxy=3, compat=3)
    b.addPairwiseBilateral(sxy=80, srgb=13, rgbims=[im0, im1, ...], compat=10)
    Qs = b.inference(10)                                  # list of (L, W_i*H_i)
"""
from numbers import Number

import numpy as np

from . import capi


class BatchedDenseCRF:
    """Batch of generic DenseCRF(N_i, L) models (features given per instance)."""

    def __init__(self, sizes, nlabels, _model=None):
        self.sizes = [int(n) for n in sizes]
        self.L = int(nlabels)
        self.model = _model if _model is not None else capi.Model.create_batch(self.sizes, self.L)
        self.offsets = np.concatenate([[0], np.cumsum(self.sizes)]).astype(np.int64)

    @property
    def N(self):
        return int(self.offsets[-1])

    def setUnaryEnergy(self, unaries):
        if len(unaries) != len(self.sizes):
            raise ValueError("need one unary per instance")
        for i, (U, n) in enumerate(zip(unaries, self.sizes)):
            U = np.asarray(U)
            if U.dtype != np.float32:
                raise ValueError("Buffer dtype mismatch, expected 'float' but got '%s'" % U.dtype)
            if U.shape != (self.L, n):
                raise ValueError("Bad shape for unary energy of instance %d (Need %s, got %s)"
                                 % (i, (self.L, n), U.shape))
            self.model.set_unary_instance(i, U)

    def addPairwiseEnergy(self, features, compat, kernel=capi.DIAG_KERNEL,
                          normalization=capi.NORMALIZE_SYMMETRIC):
        """features: one (d, N_i) float32 array per instance (same d)."""
        if len(features) != len(self.sizes):
            raise ValueError("need one feature array per instance")
        for f, n in zip(features, self.sizes):
            if np.asarray(f).ndim != 2 or np.asarray(f).shape[1] != n:
                raise ValueError("Bad shape for pairwise energy (Need (?, %d), got %s)" % (n, np.shape(f)))
        self.model.add_pairwise(np.concatenate([np.asarray(f, np.float32) for f in features], axis=1), compat,
                                kernel, normalization)

    def _split(self, Q):
        return [np.ascontiguousarray(Q[:, a:b]) for a, b in zip(self.offsets[:-1], self.offsets[1:])]

    def inference(self, n_iterations):
        return self._split(self.model.inference(n_iterations))

    def inference_device(self, n_iterations, Q_ptr, stream=None):
        """Whole batch on device: Q_ptr -> [N_total][L] float32 (instances concatenated)."""
        self.model.inference_device(n_iterations, Q_ptr, None, stream)

    def capture(self, n_iterations, Q_ptr):
        """Record the n-iteration inference into a hipGraph writing Q_ptr."""
        capi.check(capi.lib().dcrf_capture_inference(self.model.h, n_iterations, None, Q_ptr))

    def replay(self, stream=None):
        capi.check(capi.lib().dcrf_launch_graph(self.model.h, stream))


class BatchedDenseCRF2D(BatchedDenseCRF):
    """Batch of DenseCRF2D(W_i, H_i, L) models."""

    def __init__(self, shapes, nlabels):
        self.shapes = [(int(w), int(h)) for w, h in shapes]
        model = capi.Model.create_batch_2d([w for w, _ in self.shapes], [h for _, h in self.shapes], nlabels)
        super().__init__([w * h for w, h in self.shapes], nlabels, _model=model)

    def addPairwiseGaussian(self, sxy, compat, kernel=capi.DIAG_KERNEL, normalization=capi.NORMALIZE_SYMMETRIC):
        if isinstance(sxy, Number):
            sxy = (sxy, sxy)
        self.model.add_gaussian(sxy[0], sxy[1], compat, kernel, normalization)

    def addPairwiseBilateral(self, sxy, srgb, rgbims, compat, kernel=capi.DIAG_KERNEL,
                             normalization=capi.NORMALIZE_SYMMETRIC):
        if isinstance(sxy, Number):
            sxy = (sxy, sxy)
        if isinstance(srgb, Number):
            srgb = (srgb, srgb, srgb)
        if len(rgbims) != len(self.shapes):
            raise ValueError("need one image per instance")
        for im, (w, h) in zip(rgbims, self.shapes):
            if np.asarray(im).shape != (h, w, 3):
                raise ValueError("Bad shape for pairwise bilateral (Need %s, got %s)" % ((h, w, 3), np.shape(im)))
        rgb = np.concatenate([np.ascontiguousarray(im, np.uint8).ravel() for im in rgbims])
        self.model.add_bilateral(sxy[0], sxy[1], srgb[0], srgb[1], srgb[2], rgb, compat, kernel, normalization)
