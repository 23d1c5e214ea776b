# Buffer objects with the reference's MatrixXf / VectorXf semantics
# (pydensecrf/eigen.pxd, pydensecrf/eigen.pyx): a MatrixXf of shape (rows, cols) is
# stored column-major, i.e. memory [cols][rows], and exposed with strides
# (4, 4*rows), so np.array(Q) has the (L, N) shape of the reference.

cdef class VectorXf:
    cdef float* data
    cdef Py_ssize_t shape[1]
    cdef Py_ssize_t strides[1]
    cdef int exports
    cdef void set_size(self, Py_ssize_t n) except *


cdef class MatrixXf:
    cdef float* data
    cdef Py_ssize_t shape[2]
    cdef Py_ssize_t strides[2]
    cdef int exports
    cdef void set_shape(self, Py_ssize_t rows, Py_ssize_t cols) except *


cdef MatrixXf c_matrixXf(float[:, ::1] arr)
cdef VectorXf c_vectorXf(float[::1] arr)
