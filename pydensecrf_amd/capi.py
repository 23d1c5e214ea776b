"""ctypes binding of the C ABI in include/dcrf_hip.h (libdcrf_hip.so, built in-tree).

The Cython module ``pydensecrf_amd.densecrf`` is the drop-in front end; this module
gives Python direct access to the rest of the ABI: batched models, device-pointer
entry points (used with torch tensors), lattice introspection and timing.  Loading
fails loudly if the library has not been built.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdcrf_hip.so")

OK, ERR_VALUE, ERR_MEMORY, ERR_DEVICE, ERR_OVERFLOW, ERR_STATE = 0, -1, -2, -3, -4, -5
NO_NORMALIZATION, NORMALIZE_BEFORE, NORMALIZE_AFTER, NORMALIZE_SYMMETRIC = 0, 1, 2, 3
CONST_KERNEL, DIAG_KERNEL, FULL_KERNEL = 0, 1, 2
OBJECTIVE_LOG_LIKELIHOOD, OBJECTIVE_HAMMING, OBJECTIVE_IOU = 0, 1, 2
COMPAT_POTTS, COMPAT_DIAGONAL, COMPAT_MATRIX = 0, 1, 2
MAX_DIM, MAX_LABELS, MAX_TERMS = 15, 1024, 8

_P = ctypes.c_void_p
_F = ctypes.POINTER(ctypes.c_float)
_I = ctypes.POINTER(ctypes.c_int)
_D = ctypes.POINTER(ctypes.c_double)
_S = ctypes.POINTER(ctypes.c_int16)
_U8 = ctypes.POINTER(ctypes.c_uint8)
_LL = ctypes.POINTER(ctypes.c_longlong)
_c = ctypes.c_int
_f = ctypes.c_float
_D_val = ctypes.c_double

# name -> (restype, argtypes); every symbol declared in include/dcrf_hip.h
SIGNATURES = {
    "dcrf_last_error": (ctypes.c_char_p, []),
    "dcrf_device_count": (_c, []),
    "dcrf_create": (_c, [_c, _c, ctypes.POINTER(_P)]),
    "dcrf_create_2d": (_c, [_c, _c, _c, ctypes.POINTER(_P)]),
    "dcrf_create_batch": (_c, [_c, _I, _c, ctypes.POINTER(_P)]),
    "dcrf_create_batch_2d": (_c, [_c, _I, _I, _c, ctypes.POINTER(_P)]),
    "dcrf_destroy": (None, [_P]),
    "dcrf_set_unary": (_c, [_P, _F]),
    "dcrf_set_unary_instance": (_c, [_P, _c, _F]),
    "dcrf_add_pairwise": (_c, [_P, _F, _c, _c, _F, _c, _c]),
    "dcrf_add_gaussian_2d": (_c, [_P, _f, _f, _c, _F, _c, _c]),
    "dcrf_add_bilateral_2d": (_c, [_P, _f, _f, _f, _f, _f, _U8, _c, _F, _c, _c]),
    "dcrf_inference": (_c, [_P, _c, _F]),
    "dcrf_start_inference": (_c, [_P, _F]),
    "dcrf_step_inference": (_c, [_P, _F, _F, _F]),
    "dcrf_kl_divergence": (_c, [_P, _F, _D]),
    "dcrf_map": (_c, [_P, _c, _S]),
    "dcrf_current_map": (_c, [_P, _F, _S]),
    "dcrf_current_map_device": (_c, [_P, _P, _P, _P]),
    "dcrf_apply_pairwise": (_c, [_P, _c, _F, _F, _c]),
    "dcrf_inference_device": (_c, [_P, _c, _P, _P, _P]),
    "dcrf_capture_inference": (_c, [_P, _c, _P, _P]),
    "dcrf_inference_device_noalloc": (_c, [_P, _c, _P, _P, _P]),
    "dcrf_launch_graph": (_c, [_P, _P]),
    "dcrf_set_compat": (_c, [_P, _c, _F]),
    "dcrf_set_kernel_weights": (_c, [_P, _F]),
    "dcrf_num_compat_params": (_c, [_P]),
    "dcrf_get_compat_params": (_c, [_P, _F]),
    "dcrf_gradient": (_c, [_P, _c, _S, _f, _D, _F, _F, _F, _F]),
    "dcrf_num_kernel_params": (_c, [_P]),
    "dcrf_set_compat_params": (_c, [_P, _F]),
    "dcrf_release_cached_memory": (_c, [ctypes.POINTER(ctypes.c_size_t)]),
    "dcrf_unary_energy": (_c, [_P, _S, _F]),
    "dcrf_unary_from_softmax_device": (_c, [_P, _c, _c, _D_val, _D_val, _P, _P]),
    "dcrf_unary_from_labels_device": (_c, [_P, _c, _c, _D_val, _c, _P, _P]),
    "dcrf_pairwise_energy": (_c, [_P, _S, _c, _F]),
    "dcrf_gradient_objective": (_c, [_P, _c, _c, _S, _f, _F, _c, _D, _F, _F, _F, _F]),
    "dcrf_set_logistic_unary": (_c, [_P, _F, _c, _F]),
    "dcrf_num_unary_params": (_c, [_P]),
    "dcrf_get_unary_params": (_c, [_P, _F]),
    "dcrf_set_unary_params": (_c, [_P, _F]),
    "dcrf_unary_param_gradient": (_c, [_P, _F]),
    "dcrf_set_blur_pairs": (_c, [_P, _c]),
    "dcrf_set_blur_steps": (_c, [_P, _c]),
    "dcrf_set_q_blocks": (_c, [_P, _c]),
    "dcrf_set_blur_lanes": (_c, [_P, _c, _c]),
    "dcrf_get_kernel_params": (_c, [_P, _F]),
    "dcrf_set_kernel_params": (_c, [_P, _F]),
    "dcrf_forward_device": (_c, [_P, _c, _P, _P, _P, _P, _P]),
    "dcrf_backward_device": (_c, [_P, _P, _P, _P, _P, _P, _P]),
    "dcrf_num_variables": (_c, [_P]),
    "dcrf_num_labels": (_c, [_P]),
    "dcrf_num_terms": (_c, [_P]),
    "dcrf_num_instances": (_c, [_P]),
    "dcrf_term_info": (_c, [_P, _c, _I, _I]),
    "dcrf_get_lattice": (_c, [_P, _c, _c, _I, _I, _F, _I, _S, _F]),
    "dcrf_term_instance_vertices": (_c, [_P, _c, _c, _I]),
    "dcrf_set_splat_chunk": (_c, [_P, _c]),
    "dcrf_set_hash_load": (_c, [_P, _f]),
    "dcrf_set_splat_tile": (_c, [_P, _c]),
    "dcrf_set_meanfield_kernel": (_c, [_P, _c, _c]),
    "dcrf_enable_timing": (_c, [_P, _c]),
    "dcrf_kernel_stats": (_c, [_P, _c, _LL, _D, _D]),
    "dcrf_iteration_bytes": (_c, [_P, _D]),
    "dcrf_splat_layout": (_c, [_P, _I, _I, _I, _I, _I, _I]),
}

_lib = None


def lib():
    """The loaded library with typed entry points (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libdcrf_hip.so is not built; run __graft_entry__.build() or "
                              "python -m pydensecrf_amd._build")
        l = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(l, name)
            fn.restype = res
            fn.argtypes = args
        _lib = l
    return _lib


class DCRFError(RuntimeError):
    pass


def check(status):
    if status == OK:
        return
    msg = lib().dcrf_last_error().decode("utf-8", "replace")
    if status in (ERR_VALUE, ERR_OVERFLOW):
        raise ValueError(msg)
    if status == ERR_MEMORY:
        raise MemoryError(msg)
    raise DCRFError(msg)


def fptr(a):
    return a.ctypes.data_as(_F)


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def compat_args(compat, L):
    """(kind, float32 array) of a compatibility given like the reference's _labelcomp."""
    from numbers import Number
    if isinstance(compat, Number):
        return COMPAT_POTTS, np.array([compat], dtype=np.float32)
    c = np.asarray(compat)
    if c.dtype != np.float32:
        raise ValueError("Buffer dtype mismatch, expected 'float' but got '%s'" % c.dtype)
    c = np.ascontiguousarray(c)
    if c.ndim == 1:
        return COMPAT_DIAGONAL, c
    if c.ndim == 2:
        return COMPAT_MATRIX, c
    raise ValueError("LabelCompatibility of dimension >2 not meaningful.")


class Model:
    """Owner of a dcrf_model handle (plain, 2-D or batched)."""

    def __init__(self, handle, owner=True):
        self.h = handle
        self.owner = owner
        self._tuned = False

    @classmethod
    def borrow(cls, address):
        """A view of a model owned elsewhere (e.g. the drop-in module's, by its address):
        closing it does not destroy the model."""
        return cls(_P(address), owner=False)

    @classmethod
    def _created(cls, h):
        m = cls(h)
        from . import tune
        if tune.enabled():
            tune.apply_build_knobs(m)  # cached splat tile / hash load, before any term
        return m

    @classmethod
    def create(cls, N, L):
        h = _P()
        check(lib().dcrf_create(N, L, ctypes.byref(h)))
        return cls._created(h)

    @classmethod
    def create_2d(cls, W, H, L):
        h = _P()
        check(lib().dcrf_create_2d(W, H, L, ctypes.byref(h)))
        return cls._created(h)

    @classmethod
    def create_batch(cls, Ns, L):
        a = np.ascontiguousarray(Ns, dtype=np.int32)
        h = _P()
        check(lib().dcrf_create_batch(len(a), a.ctypes.data_as(_I), L, ctypes.byref(h)))
        return cls._created(h)

    @classmethod
    def create_batch_2d(cls, Ws, Hs, L):
        w = np.ascontiguousarray(Ws, dtype=np.int32)
        hh = np.ascontiguousarray(Hs, dtype=np.int32)
        h = _P()
        check(lib().dcrf_create_batch_2d(len(w), w.ctypes.data_as(_I), hh.ctypes.data_as(_I), L, ctypes.byref(h)))
        return cls._created(h)

    def tuned(self):
        """Apply the cached scheduling knobs of this model's shape (pydensecrf_amd.tune),
        once, before its first computation."""
        if not self._tuned and self.nterms:
            self._tuned = True
            from . import tune
            if tune.enabled():
                tune.apply_cached(self)
        return self

    def close(self):
        if self.h:
            if self.owner:
                lib().dcrf_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- sizes
    @property
    def N(self):
        return lib().dcrf_num_variables(self.h)

    @property
    def L(self):
        return lib().dcrf_num_labels(self.h)

    @property
    def nterms(self):
        return lib().dcrf_num_terms(self.h)

    @property
    def ninstances(self):
        return lib().dcrf_num_instances(self.h)

    # -- setup
    def set_unary(self, U_LxN):
        U = f32(U_LxN)
        if U.shape != (self.L, self.N):
            raise ValueError("unary must be %s, got %s" % ((self.L, self.N), U.shape))
        check(lib().dcrf_set_unary(self.h, fptr(U)))

    def set_unary_instance(self, i, U_LxNi):
        U = f32(U_LxNi)
        check(lib().dcrf_set_unary_instance(self.h, i, fptr(U)))

    def add_pairwise(self, feat_dxN, compat, ktype=DIAG_KERNEL, ntype=NORMALIZE_SYMMETRIC):
        f = f32(feat_dxN)
        if f.ndim != 2 or f.shape[1] != self.N:
            raise ValueError("features must be (d, %d), got %s" % (self.N, f.shape))
        kind, c = compat_args(compat, self.L)
        check(lib().dcrf_add_pairwise(self.h, fptr(f), f.shape[0], kind, fptr(c), ktype, ntype))

    def add_gaussian(self, sx, sy, compat, ktype=DIAG_KERNEL, ntype=NORMALIZE_SYMMETRIC):
        kind, c = compat_args(compat, self.L)
        check(lib().dcrf_add_gaussian_2d(self.h, sx, sy, kind, fptr(c), ktype, ntype))

    def add_bilateral(self, sx, sy, sr, sg, sb, rgb, compat, ktype=DIAG_KERNEL, ntype=NORMALIZE_SYMMETRIC):
        im = np.ascontiguousarray(rgb, dtype=np.uint8)
        if im.size != self.N * 3:
            raise ValueError("rgb must hold N*3 bytes")
        kind, c = compat_args(compat, self.L)
        check(lib().dcrf_add_bilateral_2d(self.h, sx, sy, sr, sg, sb, im.ctypes.data_as(_U8), kind, fptr(c),
                                          ktype, ntype))

    # -- inference on host buffers; Q returned as (L, N)
    def _buf(self):
        return np.empty((self.N, self.L), dtype=np.float32)

    def inference(self, n):
        self.tuned()
        Q = self._buf()
        check(lib().dcrf_inference(self.h, n, fptr(Q)))
        return Q.T

    def start(self):
        self.tuned()
        Q = self._buf()
        check(lib().dcrf_start_inference(self.h, fptr(Q)))
        return Q.T

    def step(self, Q_LxN):
        Q = np.ascontiguousarray(f32(Q_LxN).T)
        t1, t2 = self._buf(), self._buf()
        check(lib().dcrf_step_inference(self.h, fptr(Q), fptr(t1), fptr(t2)))
        return Q.T, t1.T, t2.T

    def map(self, n):
        """DenseCRF::map(n) (densecrf.cpp:132-137): labels (N,) int16 after n iterations."""
        self.tuned()
        out = np.empty(self.N, np.int16)
        check(lib().dcrf_map(self.h, n, out.ctypes.data_as(_S)))
        return out

    def current_map(self, Q_LxN):
        """DenseCRF::currentMap(Q) (densecrf.cpp:202-211): the first label of each pixel's
        largest Q, (N,) int16."""
        Q = np.ascontiguousarray(f32(Q_LxN).T)
        out = np.empty(self.N, np.int16)
        check(lib().dcrf_current_map(self.h, fptr(Q), out.ctypes.data_as(_S)))
        return out

    def current_map_device(self, Q_ptr, labels_ptr, stream=None):
        """currentMap of a device Q [N][L] float32 into device labels [N] int16."""
        check(lib().dcrf_current_map_device(self.h, Q_ptr, labels_ptr, stream))

    def kl(self, Q_LxN):
        Q = np.ascontiguousarray(f32(Q_LxN).T)
        out = ctypes.c_double()
        check(lib().dcrf_kl_divergence(self.h, fptr(Q), ctypes.byref(out)))
        return out.value

    def apply(self, term, Q_LxN, transpose=False):
        Q = np.ascontiguousarray(f32(Q_LxN).T)
        out = self._buf()
        check(lib().dcrf_apply_pairwise(self.h, term, fptr(Q), fptr(out), int(bool(transpose))))
        return out.T

    def inference_device(self, n, Q_ptr, U_ptr=None, stream=None):
        self.tuned()
        check(lib().dcrf_inference_device(self.h, n, U_ptr, Q_ptr, stream))

    def inference_device_noalloc(self, n, Q_ptr, U_ptr=None, stream=None):
        """inference_device with the vertex rows in the library's static scratch array
        (no device allocation; the model must have run inference_device once)."""
        check(lib().dcrf_inference_device_noalloc(self.h, n, U_ptr, Q_ptr, stream))

    def capture(self, n, Q_ptr, U_ptr=None):
        self.tuned()
        check(lib().dcrf_capture_inference(self.h, n, U_ptr, Q_ptr))

    def launch_graph(self, stream=None):
        check(lib().dcrf_launch_graph(self.h, stream))

    # -- parameters and gradients
    def set_compat(self, term, params):
        p = f32(np.atleast_1d(params))
        check(lib().dcrf_set_compat(self.h, term, fptr(np.ascontiguousarray(p.ravel()))))

    def set_kernel_weights(self, w):
        if w is None:
            check(lib().dcrf_set_kernel_weights(self.h, None))
            return
        a = f32(w)
        if a.size != self.nterms:
            raise ValueError("need one weight per pairwise term")
        check(lib().dcrf_set_kernel_weights(self.h, fptr(a)))

    def compat_params(self):
        p = np.empty(max(lib().dcrf_num_compat_params(self.h), 1), np.float32)
        check(lib().dcrf_get_compat_params(self.h, fptr(p)))
        return p[:lib().dcrf_num_compat_params(self.h)]

    def unary_energy(self, labels):
        """DenseCRF::unaryEnergy: U(l_i, i) per variable."""
        l = np.ascontiguousarray(labels, dtype=np.int16).ravel()
        if l.size != self.N:
            raise ValueError("need one label per variable (%d)" % self.N)
        out = np.empty(self.N, np.float32)
        check(lib().dcrf_unary_energy(self.h, l.ctypes.data_as(_S), fptr(out)))
        return out

    def pairwise_energy(self, labels, term=-1):
        """DenseCRF::pairwiseEnergy: -0.5 (message of the one-hot labelling)(l_i, i)."""
        l = np.ascontiguousarray(labels, dtype=np.int16).ravel()
        if l.size != self.N:
            raise ValueError("need one label per variable (%d)" % self.N)
        out = np.empty(self.N, np.float32)
        check(lib().dcrf_pairwise_energy(self.h, l.ctypes.data_as(_S), term, fptr(out)))
        return out

    def set_compat_params(self, p):
        a = np.ascontiguousarray(p, dtype=np.float32).ravel()
        if a.size != lib().dcrf_num_compat_params(self.h):
            raise ValueError("expected %d compatibility parameters, got %d"
                             % (lib().dcrf_num_compat_params(self.h), a.size))
        check(lib().dcrf_set_compat_params(self.h, fptr(a)))

    def gradient_objective(self, n, gt, objective=0, param=0.0, class_weight=None, kernel=False, unary_grad=True):
        """DenseCRF::gradient with LOG_LIKELIHOOD (param = robust), HAMMING (class_weight, or
        param = class_weight_pow) or IOU: (objective, unary_grad (L, N) or None when not
        asked for, compat_grad packed, kernel-weight grad [T], kernel-parameter grad or
        None)."""
        self.tuned()
        g = np.ascontiguousarray(gt, dtype=np.int16)
        if g.shape != (self.N,):
            raise ValueError("gt must have one label per variable (%d), got %s" % (self.N, g.shape))
        obj = ctypes.c_double()
        ug = self._buf() if unary_grad else None
        cg = np.empty(max(lib().dcrf_num_compat_params(self.h), 1), np.float32)
        kg = np.empty(max(self.nterms, 1), np.float32)
        nk = lib().dcrf_num_kernel_params(self.h)
        kp = np.empty(max(nk, 1), np.float32)
        cw = None if class_weight is None else np.ascontiguousarray(class_weight, dtype=np.float32).ravel()
        check(lib().dcrf_gradient_objective(self.h, n, objective, g.ctypes.data_as(_S), param,
                                            None if cw is None else fptr(cw), 0 if cw is None else cw.size,
                                            ctypes.byref(obj), fptr(ug) if unary_grad else None, fptr(cg), fptr(kg),
                                            fptr(kp) if kernel else None))
        return (obj.value, ug.T if unary_grad else None, cg[:lib().dcrf_num_compat_params(self.h)],
                kg[:self.nterms], kp[:nk] if kernel else None)

    def set_logistic_unary(self, Lmat, feat):
        """LogisticUnaryEnergy: U = Lmat feat, Lmat (L, F), feat (F, N)."""
        a = np.ascontiguousarray(Lmat, dtype=np.float32)
        f = np.ascontiguousarray(feat, dtype=np.float32)
        if a.ndim != 2 or f.ndim != 2 or a.shape[0] != self.L or f.shape != (a.shape[1], self.N):
            raise ValueError("logistic unary: need L (%d, F) and f (F, %d), got %s and %s"
                             % (self.L, self.N, a.shape, f.shape))
        check(lib().dcrf_set_logistic_unary(self.h, fptr(a), a.shape[1], fptr(f)))

    def unary_params(self):
        n = lib().dcrf_num_unary_params(self.h)
        p = np.empty(max(n, 1), np.float32)
        check(lib().dcrf_get_unary_params(self.h, fptr(p)))
        return p[:n]

    def set_unary_params(self, p):
        a = np.ascontiguousarray(p, dtype=np.float32).ravel()
        if a.size != lib().dcrf_num_unary_params(self.h):
            raise ValueError("expected %d unary parameters, got %d" % (lib().dcrf_num_unary_params(self.h), a.size))
        check(lib().dcrf_set_unary_params(self.h, fptr(a)))

    def unary_param_gradient(self):
        n = lib().dcrf_num_unary_params(self.h)
        g = np.empty(max(n, 1), np.float32)
        check(lib().dcrf_unary_param_gradient(self.h, fptr(g)))
        return g[:n]

    def set_blur_lanes(self, labels_per_lane, below_vertices=0):
        """Blur passes of fewer than below_vertices vertices (all when 0) with the given
        labels per lane (0 = default)."""
        check(lib().dcrf_set_blur_lanes(self.h, int(labels_per_lane), int(below_vertices)))

    def set_q_blocks(self, on):
        """Keep Q between inference iterations in 8 x 8 pixel blocks (opt-in)."""
        check(lib().dcrf_set_q_blocks(self.h, int(bool(on))))

    def set_blur_pairs(self, on):
        check(lib().dcrf_set_blur_pairs(self.h, 1 if on else 0))

    def set_blur_steps(self, steps):
        """Most blur steps per multi-step pass, 2 (default) or 3 (the same bits for both)."""
        check(lib().dcrf_set_blur_steps(self.h, int(steps)))

    def kernel_params(self):
        n = lib().dcrf_num_kernel_params(self.h)
        p = np.empty(max(n, 1), np.float32)
        check(lib().dcrf_get_kernel_params(self.h, fptr(p)))
        return p[:n]

    def set_kernel_params(self, p):
        a = np.ascontiguousarray(p, dtype=np.float32).ravel()
        if a.size != lib().dcrf_num_kernel_params(self.h):
            raise ValueError("expected %d kernel parameters, got %d" % (lib().dcrf_num_kernel_params(self.h), a.size))
        check(lib().dcrf_set_kernel_params(self.h, fptr(a)))

    def gradient(self, n, gt, robust=0.0, kernel=False):
        """DenseCRF::gradient with LogLikelihood: (objective, unary_grad (L, N),
        compat_grad packed, kernel-weight grad [T][, kernel-parameter grad])."""
        self.tuned()
        g = np.ascontiguousarray(gt, dtype=np.int16)
        if g.shape != (self.N,):
            raise ValueError("gt must have one label per variable (%d), got %s" % (self.N, g.shape))
        obj = ctypes.c_double()
        ug = self._buf()
        cg = np.empty(max(lib().dcrf_num_compat_params(self.h), 1), np.float32)
        kg = np.empty(max(self.nterms, 1), np.float32)
        nk = lib().dcrf_num_kernel_params(self.h)
        kp = np.empty(max(nk, 1), np.float32) if kernel else None
        check(lib().dcrf_gradient(self.h, n, g.ctypes.data_as(_S), robust, ctypes.byref(obj), fptr(ug), fptr(cg),
                                  fptr(kg), fptr(kp) if kernel else None))
        out = (obj.value, ug.T, cg[:lib().dcrf_num_compat_params(self.h)], kg[:self.nterms])
        return out + (kp[:nk],) if kernel else out

    def forward_device(self, n, U_ptr, Q_ptr, compat_ptr=None, kw_ptr=None, stream=None):
        self.tuned()
        check(lib().dcrf_forward_device(self.h, n, U_ptr, compat_ptr, kw_ptr, Q_ptr, stream))

    def backward_device(self, dQ_ptr, dU_ptr, dcompat_ptr=None, dkw_ptr=None, stream=None, dkernel_ptr=None):
        check(lib().dcrf_backward_device(self.h, dQ_ptr, dU_ptr, dcompat_ptr, dkw_ptr, dkernel_ptr, stream))

    # -- introspection
    def term_info(self, term):
        d, V = ctypes.c_int(), ctypes.c_int()
        check(lib().dcrf_term_info(self.h, term, ctypes.byref(d), ctypes.byref(V)))
        return d.value, V.value

    def instance_vertices(self, term, inst=0):
        V = ctypes.c_int()
        check(lib().dcrf_term_instance_vertices(self.h, term, inst, ctypes.byref(V)))
        return V.value

    def lattice(self, term, inst=0, Ni=None):
        """Lattice arrays of one instance: offset, rank, bary (Ni, d+1), nbr (d+1, V, 2),
        keys (V, d), norm (Ni,)."""
        d, _ = self.term_info(term)
        V = self.instance_vertices(term, inst)
        if Ni is None:
            Ni = self.N
        off = np.empty((Ni, d + 1), np.int32)
        rank = np.empty((Ni, d + 1), np.int32)
        bary = np.empty((Ni, d + 1), np.float32)
        nbr = np.empty((d + 1, V, 2), np.int32)
        keys = np.empty((V, d), np.int16)
        norm = np.empty(Ni, np.float32)
        check(lib().dcrf_get_lattice(self.h, term, inst, off.ctypes.data_as(_I), rank.ctypes.data_as(_I),
                                     fptr(bary), nbr.ctypes.data_as(_I), keys.ctypes.data_as(_S), fptr(norm)))
        return dict(offset=off, rank=rank, bary=bary, nbr=nbr, keys=keys, norm=norm, V=V, d=d)

    def set_splat_chunk(self, c):
        check(lib().dcrf_set_splat_chunk(self.h, c))

    def set_hash_load(self, load):
        check(lib().dcrf_set_hash_load(self.h, load))

    def set_splat_tile(self, side):
        check(lib().dcrf_set_splat_tile(self.h, side))

    def set_meanfield_kernel(self, mode, run_rows=0):
        """0 automatic (fused where available), 1 generic, 2 fused, 3 fused over
        LDS-staged pixel runs; run_rows: height of the pixel-run rectangles (0 = 16)."""
        check(lib().dcrf_set_meanfield_kernel(self.h, mode, run_rows))

    def enable_timing(self, on=True):
        check(lib().dcrf_enable_timing(self.h, int(bool(on))))

    KERNEL_CLASSES = ("splat", "combine", "blur", "meanfield")

    def kernel_stats(self):
        """{class: (launches, total ms, algorithmic bytes)} of the timed launches."""
        out = {}
        for i, name in enumerate(self.KERNEL_CLASSES):
            n, ms, b = ctypes.c_longlong(), ctypes.c_double(), ctypes.c_double()
            check(lib().dcrf_kernel_stats(self.h, i, ctypes.byref(n), ctypes.byref(ms), ctypes.byref(b)))
            out[name] = (n.value, ms.value, b.value)
        return out

    def splat_layout(self):
        """{tiled, tile_side, n_tiles, segments, partials, meanfield} of the filter layout
        (meanfield: 0 generic, 1 fused, 2 fused over pixel runs)."""
        v = [ctypes.c_int() for _ in range(6)]
        check(lib().dcrf_splat_layout(self.h, *[ctypes.byref(x) for x in v]))
        return dict(zip(("tiled", "tile_side", "n_tiles", "segments", "partials", "meanfield"), [x.value for x in v]))

    def iteration_bytes(self):
        b = ctypes.c_double()
        check(lib().dcrf_iteration_bytes(self.h, ctypes.byref(b)))
        return b.value
