"""Differentiable mean-field inference for PyTorch (ROCm) on top of the HIP kernels.

``CRFLayer`` wraps a model with its pairwise terms added and learns the label
compatibilities, the per-term weights and, on request, the kernel parameters (the
feature scales of DIAG_KERNEL terms, the feature matrices of FULL_KERNEL terms):

    E_{t+1} = U + sum_k C_k(w_k K~_k Q_t),   Q_t = softmax(-E_t),   Q_0 = softmax(-U)

with C_k Potts (-c I), Diagonal (diag c) or Matrix ((m + m^T)/2), as in the reference's
DenseCRF (pydensecrf/densecrf/src/densecrf.cpp:115-131; labelcompatibility.cpp).  The
backward is the hand-written adjoint of libdcrf_hip.so (dcrf_backward_device, the
reference's DenseCRF::gradient recursion, densecrf.cpp:238-297): it returns gradients
for the unary, the raw compatibility parameters, the term weights and the kernel
parameters (kernel_grad, pairwise.cpp:152-163, 202-207).  A forward whose kernel
parameters changed since the last one rebuilds the lattices first
(setKernelParameters, densecrf.cpp:348-360).  Q_0..Q_n and the messages are
checkpointed by the forward; one forward/backward pair at a time per layer.

Tensors live on the GPU in the library's layout: unary and Q are (N, L) float32,
label fastest.  Everything runs on torch's current stream.
"""
import torch

from . import capi


def _stream(t):
    return torch.cuda.current_stream(t.device).cuda_stream


def _raw_len(kind, L):
    return 1 if kind == capi.COMPAT_POTTS else (L if kind == capi.COMPAT_DIAGONAL else L * L)


class _MeanField(torch.autograd.Function):
    @staticmethod
    def forward(ctx, layer, unary, compat, kweights, kparams):
        if not (unary.is_cuda and unary.dtype == torch.float32):
            raise TypeError("unary must be a float32 CUDA tensor")
        unary = unary.contiguous()
        compat = compat.detach().contiguous()
        kweights = kweights.detach().contiguous()
        layer._sync_kernel_params(kparams)
        Q = torch.empty_like(unary)
        layer._version += 1
        layer.model.forward_device(layer.n_iter, unary.data_ptr(), Q.data_ptr(), compat.data_ptr(),
                                   kweights.data_ptr(), _stream(unary))
        ctx.layer = layer
        ctx.version = layer._version
        ctx.save_for_backward(compat, kweights)
        return Q

    @staticmethod
    def backward(ctx, dQ):
        layer = ctx.layer
        if layer._version != ctx.version:
            raise RuntimeError("CRFLayer: another forward ran since this graph was built; the mean-field "
                               "checkpoints of this pass are gone (one forward/backward pair at a time)")
        compat, kweights = ctx.saved_tensors
        dQ = dQ.contiguous()
        dU = torch.empty_like(dQ)
        dc = torch.empty_like(compat)
        dk = torch.empty_like(kweights)
        dp = None
        if ctx.needs_input_grad[4] and layer.kparams.numel():
            dp = torch.empty_like(layer.kparams)
        layer.model.backward_device(dQ.data_ptr(), dU.data_ptr(), dc.data_ptr(), dk.data_ptr(), _stream(dQ),
                                    dp.data_ptr() if dp is not None else None)
        return None, dU, dc, dk, dp


class CRFLayer(torch.nn.Module):
    """Mean-field CRF layer over a ``pydensecrf_amd.capi.Model`` with its terms added.

    compat: list with one initial compatibility per term (number / (L,) / (L, L), the
    kinds the terms were created with); kweights: initial term weights (default 1).
    kparams: initial kernel parameters, packed like the reference's kernelParameters
    (default: the model's); learn_kernel makes them a trained parameter.
    """

    def __init__(self, model, n_iter, compat, kweights=None, learn_compat=True, learn_kweights=True,
                 kparams=None, learn_kernel=False):
        super().__init__()
        self.model = model
        self.n_iter = int(n_iter)
        self.L = model.L
        T = model.nterms
        if len(compat) != T:
            raise ValueError("need one compatibility per pairwise term (%d)" % T)
        flat, self.kinds = [], []
        for c in compat:
            kind, arr = capi.compat_args(c if not torch.is_tensor(c) else c.detach().cpu().float().numpy(), self.L)
            self.kinds.append(kind)
            flat.append(torch.as_tensor(arr, dtype=torch.float32).reshape(-1))
        self.compat = torch.nn.Parameter(torch.cat(flat).cuda(), requires_grad=learn_compat)
        kw = torch.ones(T) if kweights is None else torch.as_tensor(kweights, dtype=torch.float32).reshape(T)
        self.kweights = torch.nn.Parameter(kw.cuda(), requires_grad=learn_kweights)
        kp0 = model.kernel_params()
        kp = torch.as_tensor(kp0 if kparams is None else kparams, dtype=torch.float32).reshape(-1)
        if kp.numel() != kp0.size:
            raise ValueError("need %d kernel parameters, got %d" % (kp0.size, kp.numel()))
        self.kparams = torch.nn.Parameter(kp.cuda(), requires_grad=learn_kernel)
        self._kp_model = torch.as_tensor(kp0, dtype=torch.float32)  # what the lattices were built with
        self._kp_seen = None  # (storage, in-place version) of the kernel parameters last compared
        self._version = 0

    def _sync_kernel_params(self, kp):
        """Rebuild the lattices when the kernel parameters moved.  Reading them back is a
        host round trip (a stream synchronisation), so it happens only when the tensor
        was replaced or written in place (an optimiser step) since the last forward."""
        if not kp.numel():
            return
        seen = (kp.data_ptr(), kp._version)
        if seen == self._kp_seen:
            return
        host = kp.float().cpu()
        if not torch.equal(host, self._kp_model):
            self.model.set_kernel_params(host.numpy())
            self._kp_model = host.clone()
        self._kp_seen = seen

    def compat_of(self, k):
        """Raw compatibility parameters of term k (a view of the flat parameter)."""
        off = sum(_raw_len(kd, self.L) for kd in self.kinds[:k])
        n = _raw_len(self.kinds[k], self.L)
        v = self.compat[off:off + n]
        return v.view(self.L, self.L) if self.kinds[k] == capi.COMPAT_MATRIX else v

    def forward(self, unary):
        """unary: (N, L) float32 CUDA tensor (label fastest).  Returns Q (N, L)."""
        if unary.shape != (self.model.N, self.L):
            raise ValueError("unary must be (N, L) = %s, got %s" % ((self.model.N, self.L), tuple(unary.shape)))
        return _MeanField.apply(self, unary, self.compat, self.kweights, self.kparams)
