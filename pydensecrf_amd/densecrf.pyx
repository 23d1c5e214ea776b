# cython: language_level=3
"""Drop-in for pydensecrf.densecrf on MI355X.

Same classes, methods, defaults and exceptions as the reference's Cython module
(pydensecrf/densecrf.pyx:1-140, pydensecrf/densecrf.pxd:1-84); every operation calls
the C ABI of libdcrf_hip.so (include/dcrf_hip.h), whose HIP kernels build the lattice
and run mean-field inference on the GPU.  Q and the step temporaries are
``pydensecrf_amd.eigen.MatrixXf`` buffers of shape (L, N), as in the reference.
"""
import array
from numbers import Number

from libc.stdint cimport int16_t, uint8_t

from pydensecrf_amd.eigen cimport MatrixXf, VectorXf, c_matrixXf, c_vectorXf


cdef extern from "dcrf_hip.h" nogil:
    ctypedef struct dcrf_model:
        pass
    int DCRF_OK, DCRF_ERR_VALUE, DCRF_ERR_MEMORY, DCRF_ERR_DEVICE, DCRF_ERR_OVERFLOW, DCRF_ERR_STATE
    int DCRF_COMPAT_POTTS, DCRF_COMPAT_DIAGONAL, DCRF_COMPAT_MATRIX
    const char* dcrf_last_error()
    int dcrf_create(int N, int L, dcrf_model** out)
    int dcrf_create_2d(int W, int H, int L, dcrf_model** out)
    void dcrf_destroy(dcrf_model* m)
    int dcrf_set_unary(dcrf_model* m, const float* U)
    int dcrf_add_pairwise(dcrf_model* m, const float* feat, int d, int ckind, const float* compat,
                          int ktype, int ntype)
    int dcrf_add_gaussian_2d(dcrf_model* m, float sx, float sy, int ckind, const float* compat,
                             int ktype, int ntype)
    int dcrf_add_bilateral_2d(dcrf_model* m, float sx, float sy, float sr, float sg, float sb,
                              const uint8_t* rgb, int ckind, const float* compat, int ktype,
                              int ntype)
    int dcrf_inference(dcrf_model* m, int n, float* Q)
    int dcrf_start_inference(dcrf_model* m, float* Q)
    int dcrf_step_inference(dcrf_model* m, float* Q, float* tmp1, float* tmp2)
    int dcrf_kl_divergence(dcrf_model* m, const float* Q, double* kl)
    int dcrf_map(dcrf_model* m, int n, int16_t* labels)
    int dcrf_current_map(dcrf_model* m, const float* Q, int16_t* labels)
    int dcrf_unary_energy(dcrf_model* m, const int16_t* labels, float* out)
    int dcrf_pairwise_energy(dcrf_model* m, const int16_t* labels, int term, float* out)
    int dcrf_set_logistic_unary(dcrf_model* m, const float* Lmat, int F, const float* feat)


# pydensecrf/densecrf.pxd:43-44 (pairwise.h:32-42)
cpdef enum NormalizationType:
    NO_NORMALIZATION = 0
    NORMALIZE_BEFORE = 1
    NORMALIZE_AFTER = 2
    NORMALIZE_SYMMETRIC = 3

cpdef enum KernelType:
    CONST_KERNEL = 0
    DIAG_KERNEL = 1
    FULL_KERNEL = 2


cdef int _check(int status) except -1:
    """Map a C ABI status to the exception the reference's binding would raise."""
    if status == DCRF_OK:
        return 0
    msg = dcrf_last_error().decode("utf-8", "replace")
    if status == DCRF_ERR_VALUE or status == DCRF_ERR_OVERFLOW:
        raise ValueError(msg)
    if status == DCRF_ERR_MEMORY:
        raise MemoryError(msg)
    raise RuntimeError(msg)


cdef class _Compat:
    """A label compatibility as parsed by the reference's _labelcomp."""
    cdef int kind
    cdef object values  # VectorXf / MatrixXf / None
    cdef float potts
    cdef float[::1] flat


cdef _Compat _labelcomp(compat):
    """Number -> Potts, 1-D -> Diagonal, 2-D -> Matrix (pydensecrf/densecrf.pyx:11-20).
    Non-float32 arrays raise ValueError through the buffer typing, like the reference."""
    cdef _Compat c = _Compat()
    cdef VectorXf v
    cdef MatrixXf m
    cdef Py_ssize_t r, k, n
    if isinstance(compat, Number):
        c.kind = DCRF_COMPAT_POTTS
        c.potts = compat
        c.flat = memoryview(bytearray(4)).cast('f')
        c.flat[0] = c.potts
    elif memoryview(compat).ndim == 1:
        v = c_vectorXf(compat)
        c.kind = DCRF_COMPAT_DIAGONAL
        c.values = v
        c.flat = array.array('f', memoryview(v).tolist() or [0.0])
    elif memoryview(compat).ndim == 2:
        m = c_matrixXf(compat)
        if m.shape[0] != m.shape[1]:
            raise ValueError("MatrixCompatibility needs a square matrix")
        c.kind = DCRF_COMPAT_MATRIX
        c.values = m
        n = m.shape[0]
        # the C ABI takes m row-major; the MatrixXf holds it column-major
        c.flat = memoryview(bytearray(4 * max(n * n, 1))).cast('f')
        for r in range(n):
            for k in range(n):
                c.flat[r * n + k] = m.data[k * n + r]
    else:
        raise ValueError("LabelCompatibility of dimension >2 not meaningful.")
    return c


cdef class Unary:
    """Base of the unary objects (pydensecrf/densecrf.pyx:23-37).  setUnary() takes
    ownership: the object is empty afterwards."""
    cdef object _u  # (L, N) float32 array or None

    cdef object move(self):
        u = self._u
        self._u = None
        return u


cdef class ConstUnary(Unary):
    def __cinit__(self, float[:, ::1] u not None):
        import numpy as np
        self._u = np.array(u, dtype=np.float32, copy=True)


cdef class LogisticUnary(Unary):
    """LogisticUnaryEnergy (unary.cpp:48-68): U = L f, L (labels, F), f (F, N); the
    product runs on the GPU when the model takes the unary."""
    def __cinit__(self, float[:, ::1] L not None, float[:, ::1] f not None):
        import numpy as np
        self._u = (np.array(L, dtype=np.float32, copy=True), np.array(f, dtype=np.float32, copy=True))


def _borrowed(size_t address):
    from pydensecrf_amd import capi
    return capi.Model.borrow(address)


cdef class DenseCRF:
    cdef dcrf_model* _this
    cdef int _nlabel
    cdef int _nvar
    cdef bint _tuned

    def __cinit__(self, int nvar, int nlabels, *_, **__):
        # Subclasses create their own model (pydensecrf/densecrf.pyx:51-62).
        self._this = NULL
        self._tuned = False
        if type(self) is DenseCRF:
            _check(dcrf_create(nvar, nlabels, &self._this))
            self._build_knobs()
        self._nvar = nvar
        self._nlabel = nlabels

    cdef _build_knobs(self):
        # cached build knobs of this size (pydensecrf_amd.tune), before any term
        from pydensecrf_amd import tune
        if tune.enabled():
            tune.apply_build_knobs(_borrowed(<size_t>self._this))

    cdef _tune(self):
        # cached scheduling knobs of this model's shape, once, before the first computation
        if not self._tuned:
            self._tuned = True
            _borrowed(<size_t>self._this).tuned()

    def __dealloc__(self):
        if self._this != NULL:
            dcrf_destroy(self._this)
            self._this = NULL

    def addPairwiseEnergy(self, float[:, ::1] features not None, compat, KernelType kernel=DIAG_KERNEL,
                          NormalizationType normalization=NORMALIZE_SYMMETRIC):
        if features.shape[1] != self._nvar:
            raise ValueError("Bad shape for pairwise energy (Need (?, {}), got {})".format(
                self._nvar, (features.shape[0], features.shape[1])))
        cdef _Compat c = _labelcomp(compat)
        cdef int d = features.shape[0]
        cdef const float* cp = &c.flat[0]
        cdef int kind = c.kind
        cdef int st
        with nogil:
            st = dcrf_add_pairwise(self._this, &features[0, 0], d, kind, cp, kernel, normalization)
        _check(st)

    def setUnary(self, Unary u):
        U = u.move()
        if U is None:
            raise ValueError("this unary object was already handed to a model")
        if isinstance(U, tuple):
            self._set_logistic(U[0], U[1])
        else:
            self.setUnaryEnergy(U)

    cdef _set_logistic(self, float[:, ::1] L, float[:, ::1] f):
        # LogisticUnaryEnergy(L, f) (unary.cpp:48-52): U = L f on the device
        if L.shape[0] != self._nlabel or f.shape[1] != self._nvar or f.shape[0] != L.shape[1]:
            raise ValueError("Bad shape for the logistic unary (Need L ({}, F) and f (F, {}), got {} and {})".format(
                self._nlabel, self._nvar, (L.shape[0], L.shape[1]), (f.shape[0], f.shape[1])))
        cdef int st
        cdef int F = <int>L.shape[1]
        with nogil:
            st = dcrf_set_logistic_unary(self._this, &L[0, 0], F, &f[0, 0])
        _check(st)

    def setUnaryEnergy(self, float[:, ::1] u not None, float[:, ::1] f = None):
        if u.shape[0] != self._nlabel or u.shape[1] != self._nvar:
            raise ValueError("Bad shape for unary energy (Need {}, got {})".format(
                (self._nlabel, self._nvar), (u.shape[0], u.shape[1])))
        if f is not None:
            # DenseCRF::setUnaryEnergy(L, f) (densecrf.cpp:92-94)
            self._set_logistic(u, f)
            return
        cdef int st
        with nogil:
            st = dcrf_set_unary(self._this, &u[0, 0])
        _check(st)

    def inference(self, int niter):
        self._tune()
        cdef MatrixXf Q = MatrixXf()
        Q.set_shape(self._nlabel, self._nvar)
        cdef int st
        with nogil:
            st = dcrf_inference(self._this, niter, Q.data)
        _check(st)
        return Q

    def startInference(self):
        self._tune()
        cdef MatrixXf Q = MatrixXf()
        Q.set_shape(self._nlabel, self._nvar)
        cdef int st
        with nogil:
            st = dcrf_start_inference(self._this, Q.data)
        _check(st)
        return Q, MatrixXf(), MatrixXf()

    def stepInference(self, MatrixXf Q not None, MatrixXf tmp1 not None, MatrixXf tmp2 not None):
        if Q.shape[0] != self._nlabel or Q.shape[1] != self._nvar:
            raise ValueError("Bad shape for Q (Need {}, got {})".format(
                (self._nlabel, self._nvar), (Q.shape[0], Q.shape[1])))
        tmp1.set_shape(self._nlabel, self._nvar)
        tmp2.set_shape(self._nlabel, self._nvar)
        cdef int st
        with nogil:
            st = dcrf_step_inference(self._this, Q.data, tmp1.data, tmp2.data)
        _check(st)

    def klDivergence(self, MatrixXf Q not None):
        if Q.shape[0] != self._nlabel or Q.shape[1] != self._nvar:
            raise ValueError("Bad shape for Q (Need {}, got {})".format(
                (self._nlabel, self._nvar), (Q.shape[0], Q.shape[1])))
        cdef double kl = 0
        cdef int st
        with nogil:
            st = dcrf_kl_divergence(self._this, Q.data, &kl)
        _check(st)
        return kl

    def map(self, int niter):
        """DenseCRF::map(n) of the reference's C++ interface (densecrf.cpp:132-137), which
        its Python binding does not wrap: the MAP label of every variable, an (N,) int16
        array, after niter iterations; the same labels as np.argmax(inference(niter), 0)."""
        import numpy as np
        self._tune()
        out = np.empty(self._nvar, np.int16)
        cdef int16_t[::1] o = out
        cdef int st
        with nogil:
            st = dcrf_map(self._this, niter, &o[0])
        _check(st)
        return out

    def currentMap(self, MatrixXf Q not None):
        """DenseCRF::currentMap(Q) of the reference's C++ interface (densecrf.cpp:202-211):
        the first label of each variable's largest Q, an (N,) int16 array."""
        import numpy as np
        if Q.shape[0] != self._nlabel or Q.shape[1] != self._nvar:
            raise ValueError("Bad shape for Q (Need {}, got {})".format(
                (self._nlabel, self._nvar), (Q.shape[0], Q.shape[1])))
        out = np.empty(self._nvar, np.int16)
        cdef int16_t[::1] o = out
        cdef int st
        with nogil:
            st = dcrf_current_map(self._this, Q.data, &o[0])
        _check(st)
        return out

    def unaryEnergy(self, labels):
        """DenseCRF::unaryEnergy(l) of the reference's C++ interface (densecrf.cpp:141-153):
        U(l_i, i) per variable for a labelling of N labels, an (N,) float32 array."""
        import numpy as np
        lab = np.ascontiguousarray(labels, dtype=np.int16).ravel()
        if lab.shape[0] != self._nvar:
            raise ValueError("Bad shape for labels (Need ({},), got {})".format(self._nvar, lab.shape))
        out = np.empty(self._nvar, np.float32)
        cdef int16_t[::1] l = lab
        cdef float[::1] o = out
        cdef int st
        with nogil:
            st = dcrf_unary_energy(self._this, &l[0], &o[0])
        _check(st)
        return out

    def pairwiseEnergy(self, labels, int term=-1):
        """DenseCRF::pairwiseEnergy(l, term) of the reference's C++ interface
        (densecrf.cpp:155-177): -0.5 (C K onehot(l))(l_i, i) per variable for pairwise
        term `term` (-1: the sum over all terms), an (N,) float32 array."""
        import numpy as np
        lab = np.ascontiguousarray(labels, dtype=np.int16).ravel()
        if lab.shape[0] != self._nvar:
            raise ValueError("Bad shape for labels (Need ({},), got {})".format(self._nvar, lab.shape))
        out = np.empty(self._nvar, np.float32)
        cdef int16_t[::1] l = lab
        cdef float[::1] o = out
        cdef int st
        with nogil:
            st = dcrf_pairwise_energy(self._this, &l[0], term, &o[0])
        _check(st)
        return out

    @property
    def _handle(self):
        """Address of the underlying dcrf_model (for the ctypes bindings of the C ABI)."""
        return <size_t>self._this


cdef class DenseCRF2D(DenseCRF):
    cdef int _w
    cdef int _h

    def __cinit__(self, int w, int h, int nlabels, *_, **__):
        if type(self) is DenseCRF2D:
            _check(dcrf_create_2d(w, h, nlabels, &self._this))
            self._build_knobs()
        self._w = w
        self._h = h
        self._nvar = w * h
        self._nlabel = nlabels

    def addPairwiseGaussian(self, sxy, compat, KernelType kernel=DIAG_KERNEL,
                            NormalizationType normalization=NORMALIZE_SYMMETRIC):
        if isinstance(sxy, Number):
            sxy = (sxy, sxy)
        cdef _Compat c = _labelcomp(compat)
        cdef float sx = sxy[0], sy = sxy[1]
        cdef const float* cp = &c.flat[0]
        cdef int kind = c.kind
        cdef int st
        with nogil:
            st = dcrf_add_gaussian_2d(self._this, sx, sy, kind, cp, kernel, normalization)
        _check(st)

    def addPairwiseBilateral(self, sxy, srgb, unsigned char[:, :, ::1] rgbim not None, compat,
                             KernelType kernel=DIAG_KERNEL, NormalizationType normalization=NORMALIZE_SYMMETRIC):
        if isinstance(sxy, Number):
            sxy = (sxy, sxy)
        if isinstance(srgb, Number):
            srgb = (srgb, srgb, srgb)
        if rgbim.shape[0] != self._h or rgbim.shape[1] != self._w:
            raise ValueError("Bad shape for pairwise bilateral (Need {}, got {})".format(
                (self._h, self._w, 3), rgbim.shape))
        if rgbim.shape[2] != 3:
            raise ValueError("addPairwiseBilateral only works for RGB images. For other types, use "
                             "`utils.create_pairwise_bilateral` to construct your own pairwise energy and "
                             "add it through `addPairwiseEnergy`.")
        cdef _Compat c = _labelcomp(compat)
        cdef float sx = sxy[0], sy = sxy[1], sr = srgb[0], sg = srgb[1], sb = srgb[2]
        cdef const float* cp = &c.flat[0]
        cdef int kind = c.kind
        cdef int st
        with nogil:
            st = dcrf_add_bilateral_2d(self._this, sx, sy, sr, sg, sb, &rgbim[0, 0, 0], kind, cp,
                                       kernel, normalization)
        _check(st)


# The reference exposes the enum members at module level (dcrf.DIAG_KERNEL, ...).
globals().update(NormalizationType.__members__)
globals().update(KernelType.__members__)
