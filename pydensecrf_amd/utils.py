"""Host-side numpy helpers with the reference's signatures and results
(pydensecrf/utils.py:6-162): unaries from labels or softmax outputs, and N-D
Gaussian / bilateral feature stacks for ``DenseCRF.addPairwiseEnergy``.

Feature stacks are (d, N) float32 with rows in ``meshgrid(indexing='ij')`` axis order
((z,) y, x), each coordinate divided by its axis scale, image channels after the
coordinates (pydensecrf/utils.py:89-162).  This order differs from DenseCRF2D's own
(x, y, r, g, b) features, so the two build different (equally valid) lattices.

Each coordinate row is one scaled 1-D axis vector broadcast over the grid, written
straight into the output; the divisions follow numpy's type rules for the scale given
(a Python number divides in float32, a float64 scalar in float64 and rounds once), so
the values are those of dividing the float32 grid in place.
"""
from logging import warning
from numbers import Number

import numpy as np


def _energy32(p):
    """-log(p) evaluated in double, rounded once to float32."""
    return np.float32(-np.log(p))


def unary_from_labels(labels, n_labels, gt_prob, zero_unsure=True):
    """Unary (n_labels, N) float32 of a label map for a classifier that is right with
    probability gt_prob and spreads the rest evenly (pydensecrf/utils.py:6-48).  With
    zero_unsure, label 0 means "unknown" (every class equally likely) and labels
    1..n_labels are classes 0..n_labels-1; otherwise labels are the classes."""
    if not 0 < gt_prob < 1:
        raise AssertionError("gt_prob must lie strictly between 0 and 1")
    lab = np.asarray(labels).ravel()
    if not np.issubdtype(lab.dtype, np.integer):
        raise IndexError("labels must be integers, got %s" % lab.dtype)
    cls = lab - 1 if zero_unsure else lab
    if cls.size and (cls.min() < -n_labels or cls.max() >= n_labels):
        raise IndexError("a label is outside the %d classes" % n_labels)
    cls = np.where(cls < 0, cls + n_labels, cls)  # numpy's negative-index convention
    hit = np.arange(n_labels)[:, None] == cls[None, :]
    U = np.where(hit, _energy32(gt_prob), _energy32((1.0 - gt_prob) / (n_labels - 1)))
    if zero_unsure:
        U[:, lab == 0] = _energy32(1.0 / n_labels)
    return U.astype(np.float32, copy=False)


def compute_unary(labels, M, GT_PROB=0.5):
    """Deprecated name of unary_from_labels (pydensecrf/utils.py:51-54)."""
    warning("compute_unary is deprecated; call unary_from_labels")
    return unary_from_labels(labels, M, GT_PROB)


def unary_from_softmax(sm, scale=None, clip=1e-5):
    """Negative log class probabilities: sm (n_classes, ...) -> (n_classes, N) float32
    (pydensecrf/utils.py:57-81).  scale in (0, 1] mixes the probabilities with the
    uniform distribution (weight 1 - scale); clip keeps them at or above that value
    (and at most 1) so that no log is infinite.  The log runs in sm's precision, or in
    double once mixed with the (double) uniform share."""
    n = sm.shape[0]
    p = sm
    if scale is not None:
        if not 0 < scale <= 1:
            raise AssertionError("scale must lie in (0, 1]")
        share = np.float64(1 - scale) * np.float64(1.0 / n)
        p = np.asarray(scale * p).astype(np.float64, copy=False) + share
    if clip is not None:
        p = np.minimum(np.maximum(p, clip), 1.0)
    U = np.negative(np.log(p))
    return U.reshape(n, -1).astype(np.float32)


def softmax_to_unary(sm, GT_PROB=1):
    """Deprecated name of unary_from_softmax without clipping (pydensecrf/utils.py:84-87)."""
    warning("softmax_to_unary is deprecated; call unary_from_softmax")
    return unary_from_softmax(sm, None if GT_PROB == 1 else GT_PROB, clip=None)


def _fill_coordinates(out, shape, sdims):
    """Rows 0..len(shape)-1 of out (rows, prod(shape)): the 'ij' grid coordinate of each
    axis, the first len(sdims) axes divided by their scale."""
    dims = len(shape)
    for a, n in enumerate(shape):
        axis = np.arange(n, dtype=np.float32)
        if a < len(sdims):
            axis = (axis / sdims[a]).astype(np.float32, copy=False)
        view = [1] * dims
        view[a] = n
        out[a].reshape(shape)[...] = axis.reshape(view)


def create_pairwise_gaussian(sdims, shape):
    """Position features of an N-D grid: (len(sdims), prod(shape)) float32, the 'ij'
    coordinate of every axis divided by its scale (pydensecrf/utils.py:89-111)."""
    shape = tuple(int(n) for n in shape)
    if len(sdims) > len(shape):
        raise IndexError("more scales than grid axes")
    out = np.empty((len(shape), int(np.prod(shape))), np.float32)
    _fill_coordinates(out, shape, sdims)
    return out.reshape(len(sdims), -1)


def create_pairwise_bilateral(sdims, schan, img, chdim=-1):
    """Position and channel features of an N-D image (pydensecrf/utils.py:114-162):
    rows are the coordinates of the non-channel axes ('ij' order, divided by sdims)
    followed by the channels divided by schan (one number for all channels, or one per
    channel).  chdim is img's channel axis; -1 means img has a single channel and no
    channel axis."""
    img = np.asarray(img)
    chan = img[np.newaxis] if chdim == -1 else np.moveaxis(img, chdim, 0)
    grid = chan.shape[1:]
    npix = int(np.prod(grid))
    dims = len(grid)
    if len(sdims) > dims:
        raise IndexError("more scales than grid axes")
    out = np.empty((dims + chan.shape[0], npix), np.float32)
    _fill_coordinates(out, grid, sdims)
    rows = chan.reshape(chan.shape[0], npix).astype(np.float32)
    if isinstance(schan, Number):
        rows = (rows / schan).astype(np.float32, copy=False)
    else:
        if len(schan) > rows.shape[0]:
            raise IndexError("more channel scales than channels")
        for c, sc in enumerate(schan):
            rows[c] = rows[c] / sc
    out[dims:] = rows
    return out


# ---------------------------------------------------------------- device forms
# The same helpers on torch tensors that live on the GPU, written straight into the
# (N, L) label-fastest layout that CRFLayer and Model.inference_device take.

def unary_from_softmax_device(sm, scale=None, clip=1e-5):
    """unary_from_softmax for a float32 CUDA tensor sm (n_classes, ...): returns the
    (N, n_classes) unary tensor, computed in double and rounded once to float32."""
    import torch

    from . import capi
    if not (sm.is_cuda and sm.dtype == torch.float32):
        raise TypeError("sm must be a float32 CUDA tensor")
    if scale is not None and not 0 < scale <= 1:
        raise ValueError("scale must lie in (0, 1]")
    L = sm.shape[0]
    s = sm.reshape(L, -1).contiguous()
    N = s.shape[1]
    U = torch.empty((N, L), dtype=torch.float32, device=sm.device)
    capi.check(capi.lib().dcrf_unary_from_softmax_device(s.data_ptr(), L, N, -1.0 if scale is None else float(scale),
                                                         -1.0 if clip is None else float(clip), U.data_ptr(),
                                                         torch.cuda.current_stream(sm.device).cuda_stream))
    return U


def unary_from_labels_device(labels, n_labels, gt_prob, zero_unsure=True):
    """unary_from_labels for an integer CUDA tensor of labels: returns the (N, n_labels)
    unary tensor."""
    import torch

    from . import capi
    if not labels.is_cuda:
        raise TypeError("labels must be a CUDA tensor")
    if not 0 < gt_prob < 1:
        raise ValueError("gt_prob must lie strictly between 0 and 1")
    lab = labels.reshape(-1).to(torch.int32).contiguous()
    N = lab.shape[0]
    U = torch.empty((N, n_labels), dtype=torch.float32, device=labels.device)
    capi.check(capi.lib().dcrf_unary_from_labels_device(lab.data_ptr(), N, n_labels, float(gt_prob),
                                                        1 if zero_unsure else 0, U.data_ptr(),
                                                        torch.cuda.current_stream(labels.device).cuda_stream))
    return U
