"""Python mirror of the reference's public C API (include/bksp.h).

Same names, argument meaning and return convention as the reference solvers:
``bksp_dcg(A, b, tol, max_iter, x)`` returns the 0-based loop index at
convergence or a negative code, and overwrites ``x`` (the initial guess) with
the solution. ``A`` is a ``DSpMat`` built by ``dspmat``; a "symmetric" matrix
holds one stored triangle. Every call goes through libbksp_amd.so to the GPU.
"""
import ctypes

import numpy as np

from ._lib import DSpMat, ZMat, ZSpMat, load

NOT_CONVERGED = -1
DIVERGENCE = -2
INVALID_PARAMETERS = -3
INVALID_MATRIX_TYPE = -4
BKH_DEVICE_ERROR = -5
BKH_BAD_ARGUMENT = -6


class SpMat:
    """Owns the arrays behind a ``DSpMat`` (include/bksp.h:12-21)."""

    def __init__(self, n, row_ptr, col, val, mtype="unsymmetric"):
        self.rp = np.ascontiguousarray(row_ptr, dtype=np.int32)
        self.ci = np.ascontiguousarray(col, dtype=np.int32)
        self.val = np.ascontiguousarray(val, dtype=np.float64)
        self._n = ctypes.c_int(int(n))
        self._m = ctypes.c_int(int(n))
        self._nnz = ctypes.c_int(int(self.rp[-1]))
        self._fmt = b"csr"
        self._type = mtype.encode()
        self.s = DSpMat(self.rp.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                        self.ci.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                        ctypes.pointer(self._nnz), ctypes.pointer(self._n), ctypes.pointer(self._m),
                        self.val.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                        self._fmt, self._type)
        self.n = int(n)

    @property
    def ref(self):
        return ctypes.byref(self.s)


class Mat:
    """Owns the column-major array behind a ``dmat`` (include/bksp.h:23-27)."""

    def __init__(self, a):
        self.a = np.asfortranarray(a, dtype=np.float64)
        from ._lib import DMat
        self.s = DMat(self.a.shape[0], self.a.shape[1],
                      self.a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))

    @property
    def ref(self):
        return ctypes.byref(self.s)


def dspmat(n, row_ptr, col, val, symmetric=False):
    return SpMat(n, row_ptr, col, val, "symmetric" if symmetric else "unsymmetric")


def _dp(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def bksp_dcg(A, b, tol, max_iter, x):
    """CG (reference src/cg.c:6-71)."""
    return load().bksp_dcg(A.ref, _dp(b), tol, max_iter, _dp(x))


def dcr(A, b, tol, max_iter, x):
    """Conjugate residual (reference src/cr.c:6-80); the loop index starts at 1."""
    return load().dcr(A.ref, _dp(b), tol, max_iter, _dp(x))


def dbicg(A, b, tol, max_iter, x):
    """BiCG (reference src/bicg.c:7-88)."""
    return load().dbicg(A.ref, _dp(b), tol, max_iter, _dp(x))


def dbicgstabl(A, b, l, tol, max_iter, x):
    """BiCGSTAB(l) (reference src/bicgstabl.c:13-138); returns the outer index."""
    return load().dbicgstabl(A.ref, _dp(b), l, tol, max_iter, _dp(x))


def dblcg(A, B, tol, max_iter, X):
    """Block CG (reference src/blcg.c:15-119). B, X are ``Mat`` (X overwritten)."""
    return load().dblcg(A.ref, B.ref, tol, max_iter, X.ref)


def dblbicgstabl(A, B, l, tol, max_iter, X):
    """Block BiCGSTAB(l) (reference src/blbicgstabl.c:16-252); prints the reference's
    summary. B, X are ``Mat`` (X overwritten)."""
    return load().dblbicgstabl(A.ref, B.ref, l, tol, max_iter, X.ref)


def dbicgstab(A, b, tol, max_iter, x):
    """BiCGSTAB (reference src/bicgstab.c:5-73), run as BiCGSTAB(1)."""
    return load().dbicgstab(A.ref, _dp(b), tol, max_iter, _dp(x))


def dblbicgstab(A, B, tol, max_iter, X):
    """Block BiCGSTAB (reference src/blbicgstab.c:11-124), run as block BiCGSTAB(1)."""
    return load().dblbicgstab(A.ref, B.ref, tol, max_iter, X.ref)


def dshbicg(A, b, sigma, tol, max_iter, x):
    """Shifted BiCG (reference src/shbicg.c:4-130); x has n*(len(sigma)+1) entries."""
    sig = np.ascontiguousarray(sigma, dtype=np.float64)
    return load().dshbicg(A.ref, _dp(b), _dp(sig), len(sig), tol, max_iter, _dp(x))


def dshbicgstabl(A, b, l, sigma, tol, max_iter, x):
    """Shifted BiCGSTAB(l) (reference src/shbicgstabl.c:7-285). x has n*(len(sigma)+1)
    entries: seed first, then shift k at x[(k+1)*n:]."""
    sig = np.ascontiguousarray(sigma, dtype=np.float64)
    return load().dshbicgstabl(A.ref, _dp(b), l, _dp(sig), len(sig), tol, max_iter, _dp(x))


# ------------------------------------------------------------- complex family

class ZSpMatrix:
    """Owns the arrays behind a ``zspmat`` (include/bksp.h:113-122); values complex128."""

    def __init__(self, n, row_ptr, col, val, mtype="unsymmetric"):
        self.rp = np.ascontiguousarray(row_ptr, dtype=np.int32)
        self.ci = np.ascontiguousarray(col, dtype=np.int32)
        self.val = np.ascontiguousarray(val, dtype=np.complex128)
        self._n = ctypes.c_int(int(n))
        self._m = ctypes.c_int(int(n))
        self._nnz = ctypes.c_int(int(self.rp[-1]))
        self._fmt = b"csr"
        self._type = mtype.encode()
        self.s = ZSpMat(self.rp.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                        self.ci.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                        ctypes.pointer(self._nnz), ctypes.pointer(self._n), ctypes.pointer(self._m),
                        self.val.ctypes.data, self._fmt, self._type)
        self.n = int(n)

    @property
    def ref(self):
        return ctypes.byref(self.s)


class ZMatrix:
    """Owns the column-major complex array behind a ``zmat`` (include/bksp.h:124-128)."""

    def __init__(self, a):
        self.a = np.asfortranarray(a, dtype=np.complex128)
        self.s = ZMat(self.a.shape[0], self.a.shape[1], self.a.ctypes.data)

    @property
    def ref(self):
        return ctypes.byref(self.s)


def zspmat(n, row_ptr, col, val):
    return ZSpMatrix(n, row_ptr, col, val)


def _zp(a):
    assert a.dtype == np.complex128 and a.flags["C_CONTIGUOUS"]
    return ctypes.c_void_p(a.ctypes.data)


def zbicg(A, b, tol, max_iter, x):
    """Complex BiCG (reference src/bicg.c:91-177)."""
    return load().zbicg(A.ref, _zp(b), tol, max_iter, _zp(x))


def zbicgstabl(A, b, l, tol, max_iter, x):
    """Complex BiCGSTAB(l) (reference src/bicgstabl.c:141-276); returns the outer index."""
    return load().zbicgstabl(A.ref, _zp(b), l, tol, max_iter, _zp(x))


def zshbicgstabl(A, b, l, sigma, tol, max_iter, x):
    """Complex shifted BiCGSTAB(l) (reference src/shbicgstabl.c:288-592); x has
    n*(len(sigma)+1) entries, seed first."""
    sig = np.ascontiguousarray(sigma, dtype=np.complex128)
    return load().zshbicgstabl(A.ref, _zp(b), l, _zp(sig), len(sig), tol, max_iter, _zp(x))


def zblbicgstabl(A, B, l, tol, max_iter, X):
    """Complex block BiCGSTAB(l) (reference src/blbicgstabl.c:255-486); prints the
    reference's summary. B, X are ``ZMatrix`` (X overwritten)."""
    return load().zblbicgstabl(A.ref, B.ref, l, tol, max_iter, X.ref)
