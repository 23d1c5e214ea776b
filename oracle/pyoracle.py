"""Test infrastructure only: ctypes front ends for the two CPU checkers.

* ``OracleCRF`` / ``oracle_lattice`` wrap ``oracle/liboracle.so`` (the C restatement of
  the reference's mean-field path, oracle/dcrf_oracle.c).
* ``RefCRF`` / ``ref_lattice`` wrap ``oracle/_ref/libdcrf_ref.so``, the reference's own
  C++ compiled from its sources by oracle/Makefile (present only where it was built).

Both take and return numpy arrays in the reference's Python layouts: features (d, N),
unary (L, N), Q (L, N).  Only tests/, ``__graft_entry__.smoke()`` and bench.py's CPU
baseline use this module; the product never imports it.
"""
import ctypes
import os
from numbers import Number

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libdcrf_ref.so")

NO_NORMALIZATION, NORMALIZE_BEFORE, NORMALIZE_AFTER, NORMALIZE_SYMMETRIC = 0, 1, 2, 3
CONST_KERNEL, DIAG_KERNEL, FULL_KERNEL = 0, 1, 2
LOG_LIKELIHOOD, HAMMING, IOU = 0, 1, 2
POTTS, DIAGONAL, MATRIX = 0, 1, 2

_F = ctypes.POINTER(ctypes.c_float)
_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)
_S = ctypes.POINTER(ctypes.c_short)
_U8 = ctypes.POINTER(ctypes.c_ubyte)
_P = ctypes.c_void_p


def _fp(a):
    return a.ctypes.data_as(_F)


def _dp(a):
    return a.ctypes.data_as(_D) if a is not None else None


def _c32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def compat_kind(compat, L):
    """Number -> Potts, 1-D -> Diagonal, 2-D -> Matrix (pydensecrf/densecrf.pyx:11-20)."""
    if isinstance(compat, Number):
        return POTTS, np.array([compat], dtype=np.float32)
    c = _c32(compat)
    if c.ndim == 1:
        return DIAGONAL, c
    if c.ndim == 2:
        return MATRIX, c
    raise ValueError("LabelCompatibility of dimension >2 not meaningful.")


class _Lib:
    _cache = {}

    @classmethod
    def get(cls, path):
        if path not in cls._cache:
            if not os.path.exists(path):
                raise FileNotFoundError(path)
            cls._cache[path] = ctypes.CDLL(path)
        return cls._cache[path]


# --------------------------------------------------------------------------- oracle
class _OrcLattice(ctypes.Structure):
    _fields_ = [("N", ctypes.c_int), ("npad", ctypes.c_int), ("d", ctypes.c_int), ("V", ctypes.c_int),
                ("overflow", ctypes.c_int), ("offset", _I), ("rank", _I), ("bary", _F), ("nbr", _I),
                ("keys", _S)]


def _oracle():
    lib = _Lib.get(ORACLE_SO)
    if not getattr(lib, "_typed", False):
        lib.orc_lattice_build.argtypes = [ctypes.POINTER(_OrcLattice), _F, ctypes.c_int, ctypes.c_int]
        lib.orc_lattice_free.argtypes = [ctypes.POINTER(_OrcLattice)]
        lib.orc_compute_f.argtypes = [ctypes.POINTER(_OrcLattice), _F, _F, ctypes.c_int, ctypes.c_int]
        lib.orc_compute_d.argtypes = [ctypes.POINTER(_OrcLattice), _D, _D, ctypes.c_int, ctypes.c_int]
        lib.orc_exp_packet.argtypes = [ctypes.c_float]
        lib.orc_exp_packet.restype = ctypes.c_float
        lib.orc_sum_packet.argtypes = [_F, ctypes.c_long]
        lib.orc_sum_packet.restype = ctypes.c_float
        lib.orc_crf_new.argtypes = [ctypes.c_int, ctypes.c_int]
        lib.orc_crf_new.restype = _P
        lib.orc_crf_new_2d.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int]
        lib.orc_crf_new_2d.restype = _P
        lib.orc_crf_free.argtypes = [_P]
        lib.orc_crf_num_terms.argtypes = [_P]
        lib.orc_crf_lattice.argtypes = [_P, ctypes.c_int]
        lib.orc_crf_lattice.restype = ctypes.POINTER(_OrcLattice)
        lib.orc_crf_norm.argtypes = [_P, ctypes.c_int]
        lib.orc_crf_norm.restype = _F
        lib.orc_crf_set_unary.argtypes = [_P, _F]
        lib.orc_crf_add_pairwise.argtypes = [_P, _F, ctypes.c_int, ctypes.c_int, _F, ctypes.c_int, ctypes.c_int]
        lib.orc_crf_add_gaussian_2d.argtypes = [_P, ctypes.c_float, ctypes.c_float, ctypes.c_int, _F,
                                                ctypes.c_int, ctypes.c_int]
        lib.orc_crf_add_bilateral_2d.argtypes = [_P] + [ctypes.c_float] * 5 + [_U8, ctypes.c_int, _F,
                                                                                 ctypes.c_int, ctypes.c_int]
        lib.orc_crf_inference.argtypes = [_P, ctypes.c_int, _F]
        lib.orc_crf_start.argtypes = [_P, _F]
        lib.orc_crf_step.argtypes = [_P, _F, _F, _F]
        lib.orc_crf_kl.argtypes = [_P, _F]
        lib.orc_crf_kl.restype = ctypes.c_double
        lib.orc_crf_apply.argtypes = [_P, ctypes.c_int, _F, _F, ctypes.c_int]
        lib.orc_crf_num_compat_params.argtypes = [_P]
        lib.orc_crf_get_compat_params.argtypes = [_P, _F]
        lib.orc_crf_set_compat_params.argtypes = [_P, _F]
        lib.orc_crf_gradient.argtypes = [_P, ctypes.c_int, _S, ctypes.c_float, _F, _F]
        lib.orc_crf_gradient.restype = ctypes.c_double
        lib.orc_crf_gradient_k.argtypes = [_P, ctypes.c_int, _S, ctypes.c_float, _F, _F, _F]
        lib.orc_crf_gradient_k.restype = ctypes.c_double
        lib.orc_crf_num_kernel_params.argtypes = [_P]
        lib.orc_crf_gradient_obj.argtypes = [_P, ctypes.c_int, ctypes.c_int, _S, ctypes.c_float, _F, ctypes.c_int,
                                             _F, _F, _F]
        lib.orc_crf_gradient_obj.restype = ctypes.c_double
        lib.orc_crf_get_kernel_params.argtypes = [_P, _F]
        lib.orc_crf_set_kernel_params.argtypes = [_P, _F]
        lib.orc_lattice_gradient.argtypes = [ctypes.c_void_p, _F, _F, _F, ctypes.c_int]
        lib.orc_crf_inference_fp64.argtypes = [_P, ctypes.c_int, _D]
        lib.orc_crf_ce_fp64.argtypes = [_P, ctypes.c_int, _S, _D, _D, _D]
        lib.orc_crf_ce_fp64.restype = ctypes.c_double
        lib._typed = True
    return lib


class LatticeArrays:
    """Lattice arrays in the reference's layout: offset/rank/bary (npad, d+1) for the
    hashed points (pad lanes included), nbr (d+1, V, 2), keys (V, d) or None."""

    def __init__(self, N, d, V, offset, rank, bary, nbr, keys=None, overflow=False):
        self.N, self.d, self.V = N, d, V
        self.offset, self.rank, self.bary, self.nbr, self.keys = offset, rank, bary, nbr, keys
        self.overflow = overflow


def _copy_lattice(lat):
    N, d, V, npad = lat.N, lat.d, lat.V, lat.npad
    e = npad * (d + 1)
    off = np.ctypeslib.as_array(lat.offset, shape=(e,)).copy().reshape(npad, d + 1)
    rank = np.ctypeslib.as_array(lat.rank, shape=(e,)).copy().reshape(npad, d + 1)
    bary = np.ctypeslib.as_array(lat.bary, shape=(e,)).copy().reshape(npad, d + 1)
    nbr = np.ctypeslib.as_array(lat.nbr, shape=(2 * (d + 1) * V,)).copy().reshape(d + 1, V, 2)
    keys = np.ctypeslib.as_array(lat.keys, shape=(V * d,)).copy().reshape(V, d) if V * d else np.zeros((V, d), np.int16)
    return LatticeArrays(N, d, V, off, rank, bary, nbr, keys, bool(lat.overflow))


class OracleLattice:
    """Permutohedral lattice of the oracle for features (d, N)."""

    def __init__(self, feat_dxN):
        f = _c32(feat_dxN)
        self.d, self.N = f.shape
        self._lib = _oracle()
        self._lat = _OrcLattice()
        pm = np.ascontiguousarray(f.T)
        if self._lib.orc_lattice_build(ctypes.byref(self._lat), _fp(pm), self.N, self.d) != 0:
            raise ValueError("oracle lattice build failed")
        self.arrays = _copy_lattice(self._lat)
        self.V = self.arrays.V

    def compute(self, values_LxN, reverse=False, fp64=False):
        v = np.asarray(values_LxN)
        L = v.shape[0]
        if fp64:
            pm = np.ascontiguousarray(v.T, dtype=np.float64)
            out = np.empty_like(pm)
            self._lib.orc_compute_d(ctypes.byref(self._lat), _dp(pm), _dp(out), L, int(reverse))
        else:
            pm = np.ascontiguousarray(v.T, dtype=np.float32)
            out = np.empty_like(pm)
            self._lib.orc_compute_f(ctypes.byref(self._lat), _fp(pm), _fp(out), L, int(reverse))
        return out.T

    def gradient(self, a_LxN, b_LxN):
        """Gradient of a^T K b with respect to the features, (d, N)."""
        a = np.ascontiguousarray(_c32(a_LxN).T)
        b = np.ascontiguousarray(_c32(b_LxN).T)
        df = np.zeros((self.N, self.d), np.float32)
        self._lib.orc_lattice_gradient(ctypes.byref(self._lat), _fp(df), _fp(a), _fp(b), a.shape[1])
        return df.T

    def __del__(self):
        try:
            self._lib.orc_lattice_free(ctypes.byref(self._lat))
        except Exception:
            pass


def exp_packet(x):
    lib = _oracle()
    return np.array([lib.orc_exp_packet(float(v)) for v in np.ravel(x)], dtype=np.float32).reshape(np.shape(x))


def sum_packet(x):
    a = _c32(np.ravel(x))
    return _oracle().orc_sum_packet(_fp(a), a.size)


class _CRFBase:
    """Common numpy front end.  Subclasses bind the C entry points."""

    def __init__(self, N, L, W=None, H=None):
        self.N, self.L, self.W, self.H = N, L, W, H
        self.nterms = 0

    def _pm(self, a_LxN):
        a = _c32(a_LxN)
        if a.shape != (self.L, self.N):
            raise ValueError("expected shape %s, got %s" % ((self.L, self.N), a.shape))
        return np.ascontiguousarray(a.T)

    def _out(self):
        return np.empty((self.N, self.L), dtype=np.float32)


class OracleCRF(_CRFBase):
    def __init__(self, N=None, L=None, W=None, H=None):
        self._lib = _oracle()
        if W is not None:
            super().__init__(W * H, L, W, H)
            self._h = self._lib.orc_crf_new_2d(W, H, L)
        else:
            super().__init__(N, L)
            self._h = self._lib.orc_crf_new(N, L)

    def __del__(self):
        try:
            self._lib.orc_crf_free(self._h)
        except Exception:
            pass

    def set_unary(self, U_LxN):
        self._U = self._pm(U_LxN)
        self._lib.orc_crf_set_unary(self._h, _fp(self._U))

    def add_pairwise(self, feat_dxN, compat, ktype=DIAG_KERNEL, ntype=NORMALIZE_SYMMETRIC):
        f = _c32(feat_dxN)
        kind, c = compat_kind(compat, self.L)
        pm = np.ascontiguousarray(f.T)
        if self._lib.orc_crf_add_pairwise(self._h, _fp(pm), f.shape[0], kind, _fp(c), ktype, ntype) != 0:
            raise ValueError("oracle add_pairwise failed")
        self.nterms += 1

    def add_gaussian(self, sx, sy, compat, ktype=DIAG_KERNEL, ntype=NORMALIZE_SYMMETRIC):
        kind, c = compat_kind(compat, self.L)
        self._lib.orc_crf_add_gaussian_2d(self._h, sx, sy, kind, _fp(c), ktype, ntype)
        self.nterms += 1

    def add_bilateral(self, sx, sy, sr, sg, sb, im, compat, ktype=DIAG_KERNEL, ntype=NORMALIZE_SYMMETRIC):
        kind, c = compat_kind(compat, self.L)
        im = np.ascontiguousarray(im, dtype=np.uint8)
        self._lib.orc_crf_add_bilateral_2d(self._h, sx, sy, sr, sg, sb, im.ctypes.data_as(_U8), kind, _fp(c),
                                           ktype, ntype)
        self.nterms += 1

    def lattice(self, term):
        return _copy_lattice(self._lib.orc_crf_lattice(self._h, term).contents)

    def norm(self, term):
        return np.ctypeslib.as_array(self._lib.orc_crf_norm(self._h, term), shape=(self.N,)).copy()

    def inference(self, n):
        Q = self._out()
        self._lib.orc_crf_inference(self._h, n, _fp(Q))
        return Q.T

    def inference_fp64(self, n):
        Q = np.empty((self.N, self.L), dtype=np.float64)
        self._lib.orc_crf_inference_fp64(self._h, n, _dp(Q))
        return Q.T

    def start(self):
        Q = self._out()
        self._lib.orc_crf_start(self._h, _fp(Q))
        return Q.T

    def step(self, Q_LxN):
        Q = self._pm(Q_LxN)
        t1, t2 = self._out(), self._out()
        self._lib.orc_crf_step(self._h, _fp(Q), _fp(t1), _fp(t2))
        return Q.T, t1.T, t2.T

    def kl(self, Q_LxN):
        Q = self._pm(Q_LxN)
        return self._lib.orc_crf_kl(self._h, _fp(Q))

    def apply(self, term, Q_LxN, transpose=False):
        Q = self._pm(Q_LxN)
        out = self._out()
        self._lib.orc_crf_apply(self._h, term, _fp(Q), _fp(out), int(transpose))
        return out.T

    def compat_params(self):
        p = np.empty(self._lib.orc_crf_num_compat_params(self._h), dtype=np.float32)
        self._lib.orc_crf_get_compat_params(self._h, _fp(p))
        return p

    def set_compat_params(self, p):
        p = _c32(p)
        self._lib.orc_crf_set_compat_params(self._h, _fp(p))

    def kernel_params(self):
        p = np.empty(self._lib.orc_crf_num_kernel_params(self._h), dtype=np.float32)
        self._lib.orc_crf_get_kernel_params(self._h, _fp(p))
        return p

    def set_kernel_params(self, p):
        p = _c32(p)
        if p.size != self._lib.orc_crf_num_kernel_params(self._h):
            raise ValueError("kernel parameter count mismatch")
        if self._lib.orc_crf_set_kernel_params(self._h, _fp(p)) != 0:
            raise ValueError("oracle lattice rebuild failed")

    def gradient(self, n, gt, robust=0.0, kernel=False):
        """(objective, dL/dU (L, N), compat gradient[, kernel-parameter gradient])."""
        gt = np.ascontiguousarray(gt, dtype=np.int16)
        ug = self._out()
        cg = np.empty(self._lib.orc_crf_num_compat_params(self._h), dtype=np.float32)
        if not kernel:
            r = self._lib.orc_crf_gradient(self._h, n, gt.ctypes.data_as(_S), robust, _fp(ug), _fp(cg))
            return r, ug.T, cg
        kg = np.empty(self._lib.orc_crf_num_kernel_params(self._h), dtype=np.float32)
        r = self._lib.orc_crf_gradient_k(self._h, n, gt.ctypes.data_as(_S), robust, _fp(ug), _fp(cg), _fp(kg))
        return r, ug.T, cg, kg

    def gradient_obj(self, n, gt, objective, param=0.0, class_weight=None, kernel=False):
        """DenseCRF::gradient with objective LOG_LIKELIHOOD / HAMMING / IOU: (objective,
        dL/dU (L, N), compat gradient, kernel-parameter gradient)."""
        gt = np.ascontiguousarray(gt, dtype=np.int16)
        ug = self._out()
        cg = np.empty(max(self._lib.orc_crf_num_compat_params(self._h), 1), dtype=np.float32)
        kg = np.empty(max(self._lib.orc_crf_num_kernel_params(self._h), 1), dtype=np.float32)
        cw = None if class_weight is None else _c32(class_weight)
        r = self._lib.orc_crf_gradient_obj(self._h, n, objective, gt.ctypes.data_as(_S), param,
                                           None if cw is None else _fp(cw), 0 if cw is None else cw.size,
                                           _fp(ug), _fp(cg), _fp(kg) if kernel else None)
        if np.isnan(r) and cw is not None:
            raise ValueError("a ground-truth label has no class weight")
        return (r, ug.T, cg[:self._lib.orc_crf_num_compat_params(self._h)],
                kg[:self._lib.orc_crf_num_kernel_params(self._h)] if kernel else None)

    def ce_fp64(self, n, gt, U_LxN=None, compat_raw=None, kweights=None):
        gt = np.ascontiguousarray(gt, dtype=np.int16)
        U = None if U_LxN is None else np.ascontiguousarray(np.asarray(U_LxN, dtype=np.float64).T)
        cr = None if compat_raw is None else np.ascontiguousarray(compat_raw, dtype=np.float64)
        kw = None if kweights is None else np.ascontiguousarray(kweights, dtype=np.float64)
        return self._lib.orc_crf_ce_fp64(self._h, n, gt.ctypes.data_as(_S), _dp(U), _dp(cr), _dp(kw))


# --------------------------------------------------------------------------- reference
def ref_available():
    return os.path.exists(REF_SO)


def _ref():
    lib = _Lib.get(REF_SO)
    if not getattr(lib, "_typed", False):
        lib.ref_lattice_new.argtypes = [_F, ctypes.c_int, ctypes.c_int]
        lib.ref_lattice_new.restype = _P
        lib.ref_lattice_free.argtypes = [_P]
        lib.ref_lattice_vertices.argtypes = [_P]
        lib.ref_lattice_entries.argtypes = [_P]
        lib.ref_lattice_get.argtypes = [_P, _I, _I, _F, _I]
        lib.ref_lattice_compute.argtypes = [_P, _F, _F, ctypes.c_int, ctypes.c_int]
        lib.ref_lattice_gradient.argtypes = [_P, _F, _F, _F, ctypes.c_int]
        lib.ref_crf_new.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int]
        lib.ref_crf_new.restype = _P
        lib.ref_crf_free.argtypes = [_P]
        lib.ref_crf_set_unary.argtypes = [_P, _F]
        lib.ref_crf_set_unary_identity.argtypes = [_P, _F]
        lib.ref_crf_add_pairwise.argtypes = [_P, _F, ctypes.c_int, ctypes.c_int, _F, ctypes.c_int, ctypes.c_int]
        lib.ref_crf_add_gaussian.argtypes = [_P, ctypes.c_float, ctypes.c_float, ctypes.c_int, _F, ctypes.c_int,
                                             ctypes.c_int]
        lib.ref_crf_add_bilateral.argtypes = [_P] + [ctypes.c_float] * 5 + [_U8, ctypes.c_int, _F, ctypes.c_int,
                                                                            ctypes.c_int]
        lib.ref_crf_inference.argtypes = [_P, ctypes.c_int, _F]
        lib.ref_crf_start.argtypes = [_P, _F]
        lib.ref_crf_step.argtypes = [_P, _F, _F, _F]
        lib.ref_crf_kl.argtypes = [_P, _F]
        lib.ref_crf_kl.restype = ctypes.c_double
        lib.ref_crf_apply.argtypes = [_P, ctypes.c_int, _F, _F, ctypes.c_int]
        lib.ref_crf_num_compat_params.argtypes = [_P]
        lib.ref_crf_num_kernel_params.argtypes = [_P]
        lib.ref_crf_get_compat_params.argtypes = [_P, _F]
        lib.ref_crf_set_compat_params.argtypes = [_P, _F, ctypes.c_int]
        lib.ref_crf_get_kernel_params.argtypes = [_P, _F]
        lib.ref_crf_set_kernel_params.argtypes = [_P, _F, ctypes.c_int]
        lib.ref_crf_gradient.argtypes = [_P, ctypes.c_int, _S, ctypes.c_float, _F, _F, _F]
        lib.ref_crf_gradient.restype = ctypes.c_double
        lib.ref_crf_gradient_obj.argtypes = [_P, ctypes.c_int, ctypes.c_int, _S, ctypes.c_float, _F, ctypes.c_int,
                                             _F, _F, _F]
        lib.ref_crf_gradient_obj.restype = ctypes.c_double
        lib.ref_crf_set_unary_logistic.argtypes = [_P, _F, ctypes.c_int, _F]
        lib.ref_crf_unary_energy.argtypes = [_P, _S, _F]
        lib.ref_crf_pairwise_energy.argtypes = [_P, _S, ctypes.c_int, _F]
        lib.ref_crf_num_unary_params.argtypes = [_P]
        lib.ref_crf_get_unary_params.argtypes = [_P, _F]
        lib.ref_crf_set_unary_params.argtypes = [_P, _F, ctypes.c_int]
        lib.ref_minimize_lbfgs.argtypes = [ctypes.c_int, _F, _ENERGY_CB, _P, ctypes.c_int, _F]
        lib._typed = True
    return lib


# double energy(const float* x, float* g, int n, void* ctx)
_ENERGY_CB = ctypes.CFUNCTYPE(ctypes.c_double, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float),
                              ctypes.c_int, ctypes.c_void_p)


def ref_minimize_lbfgs(x0, fun, restart=0):
    """The reference's minimizeLBFGS (optimization.cpp:68-103, liblbfgs 1.10) from x0 on
    fun(x float32 (n,)) -> (value, gradient (n,)).  Returns the float32 result."""
    x0 = _c32(x0).ravel()
    n = x0.size

    def cb(xp, gp, nn, _ctx):
        x = np.ctypeslib.as_array(xp, shape=(nn,)).copy()
        f, g = fun(x)
        np.ctypeslib.as_array(gp, shape=(nn,))[:] = np.asarray(g, np.float32)
        return float(f)

    c = _ENERGY_CB(cb)
    out = np.empty(n, np.float32)
    _ref().ref_minimize_lbfgs(n, _fp(x0), c, None, int(restart), _fp(out))
    return out


class RefLattice:
    """The reference's Permutohedral for features (d, N) (built library, not imported source)."""

    def __init__(self, feat_dxN):
        f = _c32(feat_dxN)
        self.d, self.N = f.shape
        self._lib = _ref()
        self._h = self._lib.ref_lattice_new(_fp(f), self.d, self.N)
        self.V = self._lib.ref_lattice_vertices(self._h)
        e = self._lib.ref_lattice_entries(self._h)
        off = np.empty(e, np.int32)
        rank = np.empty(e, np.int32)
        bary = np.empty(e, np.float32)
        nbr = np.empty(2 * (self.d + 1) * self.V, np.int32)
        self._lib.ref_lattice_get(self._h, off.ctypes.data_as(_I), rank.ctypes.data_as(_I), _fp(bary),
                                  nbr.ctypes.data_as(_I))
        npad = (self.N + 3) // 4 * 4
        d1 = self.d + 1
        self.arrays = LatticeArrays(self.N, self.d, self.V, off[:npad * d1].reshape(npad, d1),
                                    rank[:npad * d1].reshape(npad, d1), bary[:npad * d1].reshape(npad, d1),
                                    nbr.reshape(d1, self.V, 2))

    def compute(self, values_LxN, reverse=False):
        v = _c32(values_LxN)
        L = v.shape[0]
        pm = np.ascontiguousarray(v.T)
        out = np.empty_like(pm)
        self._lib.ref_lattice_compute(self._h, _fp(pm), _fp(out), L, int(reverse))
        return out.T

    def gradient(self, a_LxN, b_LxN):
        a = np.ascontiguousarray(_c32(a_LxN).T)
        b = np.ascontiguousarray(_c32(b_LxN).T)
        df = np.zeros((self.N, self.d), np.float32)
        self._lib.ref_lattice_gradient(self._h, _fp(df), _fp(a), _fp(b), a.shape[1])
        return df.T

    def __del__(self):
        try:
            self._lib.ref_lattice_free(self._h)
        except Exception:
            pass


class RefCRF(_CRFBase):
    def __init__(self, N=None, L=None, W=None, H=None, identity_unary=False):
        self._lib = _ref()
        if W is None:
            W, H = N, 1
        super().__init__(W * H, L, W, H)
        self._h = self._lib.ref_crf_new(W, H, L)
        self._identity = identity_unary

    def __del__(self):
        try:
            self._lib.ref_crf_free(self._h)
        except Exception:
            pass

    def set_unary(self, U_LxN):
        U = _c32(U_LxN)
        if self._identity:
            self._lib.ref_crf_set_unary_identity(self._h, _fp(U))
        else:
            self._lib.ref_crf_set_unary(self._h, _fp(U))

    def add_pairwise(self, feat_dxN, compat, ktype=DIAG_KERNEL, ntype=NORMALIZE_SYMMETRIC):
        f = _c32(feat_dxN)
        kind, c = compat_kind(compat, self.L)
        self._lib.ref_crf_add_pairwise(self._h, _fp(f), f.shape[0], kind, _fp(c), ktype, ntype)
        self.nterms += 1

    def add_gaussian(self, sx, sy, compat, ktype=DIAG_KERNEL, ntype=NORMALIZE_SYMMETRIC):
        kind, c = compat_kind(compat, self.L)
        self._lib.ref_crf_add_gaussian(self._h, sx, sy, kind, _fp(c), ktype, ntype)
        self.nterms += 1

    def add_bilateral(self, sx, sy, sr, sg, sb, im, compat, ktype=DIAG_KERNEL, ntype=NORMALIZE_SYMMETRIC):
        kind, c = compat_kind(compat, self.L)
        im = np.ascontiguousarray(im, dtype=np.uint8)
        self._lib.ref_crf_add_bilateral(self._h, sx, sy, sr, sg, sb, im.ctypes.data_as(_U8), kind, _fp(c), ktype,
                                        ntype)
        self.nterms += 1

    def inference(self, n):
        Q = self._out()
        self._lib.ref_crf_inference(self._h, n, _fp(Q))
        return Q.T

    def start(self):
        Q = self._out()
        self._lib.ref_crf_start(self._h, _fp(Q))
        return Q.T

    def step(self, Q_LxN):
        Q = self._pm(Q_LxN)
        t1, t2 = self._out(), self._out()
        self._lib.ref_crf_step(self._h, _fp(Q), _fp(t1), _fp(t2))
        return Q.T, t1.T, t2.T

    def kl(self, Q_LxN):
        Q = self._pm(Q_LxN)
        return self._lib.ref_crf_kl(self._h, _fp(Q))

    def apply(self, term, Q_LxN, transpose=False):
        Q = self._pm(Q_LxN)
        out = self._out()
        self._lib.ref_crf_apply(self._h, term, _fp(Q), _fp(out), int(transpose))
        return out.T

    def compat_params(self):
        p = np.empty(self._lib.ref_crf_num_compat_params(self._h), dtype=np.float32)
        self._lib.ref_crf_get_compat_params(self._h, _fp(p))
        return p

    def set_compat_params(self, p):
        p = _c32(p)
        self._lib.ref_crf_set_compat_params(self._h, _fp(p), p.size)

    def kernel_params(self):
        p = np.empty(self._lib.ref_crf_num_kernel_params(self._h), dtype=np.float32)
        self._lib.ref_crf_get_kernel_params(self._h, _fp(p))
        return p

    def set_kernel_params(self, p):
        p = _c32(p)
        self._lib.ref_crf_set_kernel_params(self._h, _fp(p), p.size)

    def unary_energy(self, labels):
        l = np.ascontiguousarray(labels, dtype=np.int16)
        out = np.empty(self.N, np.float32)
        self._lib.ref_crf_unary_energy(self._h, l.ctypes.data_as(_S), _fp(out))
        return out

    def pairwise_energy(self, labels, term=-1):
        l = np.ascontiguousarray(labels, dtype=np.int16)
        out = np.empty(self.N, np.float32)
        self._lib.ref_crf_pairwise_energy(self._h, l.ctypes.data_as(_S), term, _fp(out))
        return out

    def set_logistic_unary(self, Lm, f):
        """DenseCRF::setUnaryEnergy(L, f): L (M, F), f (F, N)."""
        Lm, f = _c32(Lm), _c32(f)
        self._lib.ref_crf_set_unary_logistic(self._h, _fp(Lm), Lm.shape[1], _fp(f))

    def unary_params(self):
        p = np.empty(max(self._lib.ref_crf_num_unary_params(self._h), 1), dtype=np.float32)
        self._lib.ref_crf_get_unary_params(self._h, _fp(p))
        return p[:self._lib.ref_crf_num_unary_params(self._h)]

    def set_unary_params(self, p):
        p = _c32(p)
        self._lib.ref_crf_set_unary_params(self._h, _fp(p), p.size)

    def gradient_obj(self, n, gt, objective, param=0.0, class_weight=None, kernel=False, n_unary=None):
        """DenseCRF::gradient with objective LOG_LIKELIHOOD / HAMMING / IOU.  The unary
        gradient is the unary's own: (L, N) for the identity unary, else n_unary values."""
        gt = np.ascontiguousarray(gt, dtype=np.int16)
        ug = self._out() if n_unary is None else np.empty(max(n_unary, 1), np.float32)
        cg = np.empty(max(self._lib.ref_crf_num_compat_params(self._h), 1), dtype=np.float32)
        kg = np.empty(max(self._lib.ref_crf_num_kernel_params(self._h), 1), dtype=np.float32)
        cw = None if class_weight is None else _c32(class_weight)
        r = self._lib.ref_crf_gradient_obj(self._h, n, objective, gt.ctypes.data_as(_S), param,
                                           None if cw is None else _fp(cw), 0 if cw is None else cw.size,
                                           _fp(ug), _fp(cg), _fp(kg) if kernel else None)
        return (r, ug.T if n_unary is None else ug[:n_unary], cg[:self._lib.ref_crf_num_compat_params(self._h)],
                kg[:self._lib.ref_crf_num_kernel_params(self._h)] if kernel else None)

    def gradient(self, n, gt, robust=0.0, kernel=False):
        """(objective, dL/dU (L, N), compat gradient[, kernel-parameter gradient])."""
        gt = np.ascontiguousarray(gt, dtype=np.int16)
        ug = self._out()
        cg = np.empty(self._lib.ref_crf_num_compat_params(self._h), dtype=np.float32)
        if not kernel:
            r = self._lib.ref_crf_gradient(self._h, n, gt.ctypes.data_as(_S), robust, _fp(ug), _fp(cg), None)
            return r, ug.T, cg
        kg = np.empty(self._lib.ref_crf_num_kernel_params(self._h), dtype=np.float32)
        r = self._lib.ref_crf_gradient(self._h, n, gt.ctypes.data_as(_S), robust, _fp(ug), _fp(cg), _fp(kg))
        return r, ug.T, cg, kg
