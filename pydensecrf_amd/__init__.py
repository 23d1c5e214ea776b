"""Dense-CRF mean-field inference on MI355X (gfx950).

Modules
  densecrf  drop-in for pydensecrf.densecrf (Cython front end over libdcrf_hip.so)
  eigen     MatrixXf / VectorXf buffer objects of the reference's binding
  utils     the reference's numpy helpers (unaries, N-D pairwise features)
  capi      ctypes binding of the whole C ABI (include/dcrf_hip.h)
  batched   many images / volumes of uneven size in one model
  autograd  torch.autograd layer over the adjoint kernels
  synthetic seeded benchmark workloads

One HIP runtime per process: PyTorch's ROCm wheels carry their own HIP runtime
(torch/lib/libamdhip64.so).  If libdcrf_hip.so were loaded first it would bring in the
system runtime, and a second runtime in the same process sees no GPU.  When torch is
installed it is therefore imported before the library, which then binds to torch's
runtime (same soname).  Set PYDENSECRF_AMD_NO_TORCH=1 to skip this.
"""
import importlib.util as _ilu
import os as _os

if not _os.environ.get("PYDENSECRF_AMD_NO_TORCH") and _ilu.find_spec("torch") is not None:
    import torch as _torch  # noqa: F401
