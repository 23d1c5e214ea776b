# cython: language_level=3, boundscheck=False, wraparound=False
"""Host buffer objects with the reference's MatrixXf / VectorXf semantics.

Mirrors pydensecrf/eigen.pyx:1-95: ``matrixXf(arr)`` copies a C-contiguous float32
(rows, cols) array into column-major storage (as pydensecrf/eigen_impl.cpp:21-25 does),
and the object exposes that storage through the buffer protocol with shape
(rows, cols) and strides (4, 4*rows).  The same objects carry Q and the step
temporaries of ``DenseCRF`` (memory [N][L], the device layout of the library).
"""
from cpython.buffer cimport PyBUF_FORMAT
from libc.stdlib cimport malloc, free, realloc
from libc.string cimport memcpy
from posix.mman cimport madvise, MADV_HUGEPAGE
from posix.stdlib cimport posix_memalign

# Storage of 4 MiB and more is 2 MiB aligned and advised for transparent huge pages (as
# numpy does for its large arrays): a fresh 66 MB Q then takes ~33 page faults on first
# touch instead of ~16,000.
cdef size_t _HUGE = 4 << 20
cdef size_t _HUGE_ALIGN = 2 << 20


cdef float* _resize(float* old, size_t old_bytes, size_t nbytes) except NULL:
    """realloc's contract (contents kept up to the smaller size; old freed), with large
    blocks huge-page aligned and advised."""
    cdef void* p = NULL
    if nbytes < _HUGE:
        p = realloc(old, nbytes)
        if p == NULL:
            raise MemoryError()
        return <float*>p
    if posix_memalign(&p, _HUGE_ALIGN, nbytes) != 0:
        raise MemoryError()
    madvise(p, nbytes, MADV_HUGEPAGE)  # advice only: its result does not matter
    if old != NULL:
        memcpy(p, old, min(old_bytes, nbytes))
        free(old)
    return <float*>p


cdef class VectorXf:

    def __cinit__(self):
        self.data = NULL
        self.shape[0] = 0
        self.strides[0] = sizeof(float)
        self.exports = 0

    def __dealloc__(self):
        free(self.data)

    cdef void set_size(self, Py_ssize_t n) except *:
        cdef float* p
        if n == self.shape[0] and (self.data != NULL or n == 0):
            return
        if self.exports:
            raise BufferError("cannot resize a VectorXf while its buffer is exported")
        p = _resize(self.data, self.shape[0] * sizeof(float), max(n, 1) * sizeof(float))
        self.data = p
        self.shape[0] = n

    def __getbuffer__(self, Py_buffer* buf, int flags):
        buf.buf = <char*>self.data
        buf.format = 'f'
        buf.internal = NULL
        buf.itemsize = sizeof(float)
        buf.len = self.shape[0] * sizeof(float)
        buf.ndim = 1
        buf.obj = self
        buf.readonly = 0
        buf.shape = self.shape
        buf.strides = self.strides
        buf.suboffsets = NULL
        self.exports += 1

    def __releasebuffer__(self, Py_buffer* buf):
        self.exports -= 1


cdef class MatrixXf:

    def __cinit__(self):
        self.data = NULL
        self.shape[0] = 0
        self.shape[1] = 0
        self.strides[0] = sizeof(float)
        self.strides[1] = 0
        self.exports = 0

    def __dealloc__(self):
        free(self.data)

    cdef void set_shape(self, Py_ssize_t rows, Py_ssize_t cols) except *:
        cdef float* p
        if rows == self.shape[0] and cols == self.shape[1] and (self.data != NULL or rows * cols == 0):
            return
        if self.exports:
            raise BufferError("cannot resize a MatrixXf while its buffer is exported")
        p = _resize(self.data, self.shape[0] * self.shape[1] * sizeof(float), max(rows * cols, 1) * sizeof(float))
        self.data = p
        self.shape[0] = rows
        self.shape[1] = cols
        self.strides[0] = sizeof(float)
        self.strides[1] = rows * sizeof(float)

    def __getbuffer__(self, Py_buffer* buf, int flags):
        buf.buf = <char*>self.data
        buf.format = 'f'
        buf.internal = NULL
        buf.itemsize = sizeof(float)
        buf.len = self.shape[0] * self.shape[1] * sizeof(float)
        buf.ndim = 2
        buf.obj = self
        buf.readonly = 0
        buf.shape = self.shape
        buf.strides = self.strides
        buf.suboffsets = NULL
        self.exports += 1

    def __releasebuffer__(self, Py_buffer* buf):
        self.exports -= 1


cdef MatrixXf c_matrixXf(float[:, ::1] arr):
    """Copy a C-contiguous (rows, cols) float32 array into column-major storage."""
    cdef MatrixXf m = MatrixXf()
    cdef Py_ssize_t r, c, rows = arr.shape[0], cols = arr.shape[1]
    m.set_shape(rows, cols)
    for c in range(cols):
        for r in range(rows):
            m.data[c * rows + r] = arr[r, c]
    return m


cdef VectorXf c_vectorXf(float[::1] arr):
    cdef VectorXf v = VectorXf()
    v.set_size(arr.shape[0])
    if arr.shape[0]:
        memcpy(v.data, &arr[0], arr.shape[0] * sizeof(float))
    return v


def vectorXf(float[::1] arr not None):
    """pydensecrf/eigen.pyx:13-14."""
    return c_vectorXf(arr)


def matrixXf(float[:, ::1] arr not None):
    """pydensecrf/eigen.pyx:60-61."""
    return c_matrixXf(arr)
