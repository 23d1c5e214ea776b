"""Host-side numpy helpers with the reference's signatures and results
(pydensecrf/utils.py:6-162): unaries from labels or softmax outputs, and N-D
Gaussian / bilateral feature stacks for ``DenseCRF.addPairwiseEnergy``.

Feature stacks use ``meshgrid(indexing='ij')`` axis order ((z,) y, x), each axis divided
by its scale in float32, image channels appended after the coordinates
(pydensecrf/utils.py:89-162).  Note that this order differs from DenseCRF2D's own
(x, y, r, g, b) features, so the two build different (equally valid) lattices.
"""
from logging import warning
from numbers import Number

import numpy as np


def unary_from_labels(labels, n_labels, gt_prob, zero_unsure=True):
    """Unary from a label map with a classifier that is right with probability gt_prob
    (pydensecrf/utils.py:6-48).  With zero_unsure, label 0 means "unknown" (uniform)
    and labels 1..n_labels map to classes 0..n_labels-1."""
    assert 0 < gt_prob < 1, "`gt_prob must be in (0,1)."
    labels = np.asarray(labels).flatten()
    n_energy = -np.log((1.0 - gt_prob) / (n_labels - 1))
    p_energy = -np.log(gt_prob)
    U = np.full((n_labels, len(labels)), n_energy, dtype="float32")
    U[labels - 1 if zero_unsure else labels, np.arange(U.shape[1])] = p_energy
    if zero_unsure:
        U[:, labels == 0] = -np.log(1.0 / n_labels)
    return U


def compute_unary(labels, M, GT_PROB=0.5):
    """Deprecated alias of unary_from_labels (pydensecrf/utils.py:51-54)."""
    warning("pydensecrf.compute_unary is deprecated, use unary_from_labels instead.")
    return unary_from_labels(labels, M, GT_PROB)


def unary_from_softmax(sm, scale=None, clip=1e-5):
    """Negative log of class probabilities, shape (n_classes, ...) -> (n_classes, N)
    (pydensecrf/utils.py:57-81).  scale blends with the uniform distribution; clip
    bounds probabilities away from 0."""
    num_cls = sm.shape[0]
    if scale is not None:
        assert 0 < scale <= 1, "`scale` needs to be in (0,1]"
        uniform = np.ones(sm.shape) / num_cls
        sm = scale * sm + (1 - scale) * uniform
    if clip is not None:
        sm = np.clip(sm, clip, 1.0)
    return -np.log(sm).reshape([num_cls, -1]).astype(np.float32)


def softmax_to_unary(sm, GT_PROB=1):
    """Deprecated alias of unary_from_softmax (pydensecrf/utils.py:84-87)."""
    warning("pydensecrf.softmax_to_unary is deprecated, use unary_from_softmax instead.")
    scale = None if GT_PROB == 1 else GT_PROB
    return unary_from_softmax(sm, scale, clip=None)


def _scaled_mesh(shape, sdims):
    mesh = np.array(np.meshgrid(*[np.arange(s) for s in shape], indexing="ij"), dtype=np.float32)
    for i, s in enumerate(sdims):
        mesh[i] /= s
    return mesh


def create_pairwise_gaussian(sdims, shape):
    """Position features of an N-D grid, (len(shape), prod(shape)) float32
    (pydensecrf/utils.py:89-111)."""
    mesh = _scaled_mesh(shape, sdims)
    return mesh.reshape([len(sdims), -1])


def create_pairwise_bilateral(sdims, schan, img, chdim=-1):
    """Position plus channel features of an N-D image (pydensecrf/utils.py:114-162).
    chdim is the channel axis of img, or -1 for a single-channel image."""
    if chdim == -1:
        im_feat = img[np.newaxis].astype(np.float32)
    else:
        im_feat = np.rollaxis(img, chdim).astype(np.float32)
    if isinstance(schan, Number):
        im_feat /= schan
    else:
        for i, s in enumerate(schan):
            im_feat[i] /= s
    mesh = _scaled_mesh(im_feat.shape[1:], sdims)
    feats = np.concatenate([mesh, im_feat])
    return feats.reshape([feats.shape[0], -1])


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
        raise ValueError("`scale` needs to be in (0,1]")
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
        raise ValueError("`gt_prob must be in (0,1).")
    lab = labels.reshape(-1).to(torch.int32).contiguous()
    N = lab.shape[0]
    U = torch.empty((N, n_labels), dtype=torch.float32, device=labels.device)
    capi.check(capi.lib().dcrf_unary_from_labels_device(lab.data_ptr(), N, n_labels, float(gt_prob),
                                                        1 if zero_unsure else 0, U.data_ptr(),
                                                        torch.cuda.current_stream(labels.device).cuda_stream))
    return U
