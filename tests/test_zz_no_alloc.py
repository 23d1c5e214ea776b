"""The allocation-free inference entry point (dcrf_inference_device_noalloc) against
dcrf_inference_device, through the C ABI: bit-identical Q on two calls in a row, the
second on a smaller model (the shared scratch array keeps the first call's rows), and
again on the first model; its state and size errors."""
import numpy as np
import pytest

from pydensecrf_amd import capi
from tests.helpers import image_case

pytestmark = pytest.mark.gpu


def _model(W, H, L, seed):
    img, U, _ = image_case(W, H, L, seed=seed, sites=6)
    m = capi.Model.create_2d(W, H, L)
    m.set_unary(U)
    m.add_gaussian(3.0, 3.0, 3.0)
    m.add_bilateral(40.0, 40.0, 13.0, 13.0, 13.0, img, 10.0)
    return m


def _run(m, n, noalloc):
    import torch
    Q = torch.full((m.N * m.L,), -1.0, device="cuda", dtype=torch.float32)
    stream = torch.cuda.current_stream().cuda_stream
    if noalloc:
        m.inference_device_noalloc(n, Q.data_ptr(), None, stream)
    else:
        m.inference_device(n, Q.data_ptr(), None, stream)
    torch.cuda.synchronize()
    return Q.cpu().numpy()


def test_noalloc_matches_inference_device():
    import torch
    big = _model(160, 120, 21, 1)
    small = _model(50, 37, 5, 2)
    with pytest.raises(capi.DCRFError, match="layout is not built"):
        big.inference_device_noalloc(3, torch.empty(big.N * big.L, device="cuda").data_ptr())
    ref_big = _run(big, 5, False)
    ref_small = _run(small, 4, False)
    got_big = _run(big, 5, True)
    got_small = _run(small, 4, True)  # second call, smaller model
    assert np.array_equal(got_big.view(np.int32), ref_big.view(np.int32))
    assert np.array_equal(got_small.view(np.int32), ref_small.view(np.int32))
    again = _run(big, 5, True)  # the array holds the small model's rows now
    assert np.array_equal(again.view(np.int32), ref_big.view(np.int32))
    # the model's own buffers are untouched: the plain entry point still agrees
    assert np.array_equal(_run(big, 5, False).view(np.int32), ref_big.view(np.int32))


def test_noalloc_rejects_models_over_the_array():
    import torch
    m = _model(2048, 1536, 21, 3)  # ~130 MB of vertex rows
    _run(m, 0, False)  # layout built
    with pytest.raises(ValueError, match="allocation-free inference holds at most"):
        m.inference_device_noalloc(2, torch.empty(m.N * m.L, device="cuda").data_ptr())
