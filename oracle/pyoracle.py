"""ctypes wrapper of the CPU oracle (oracle/liboracle.so) -- test infrastructure.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
this module; the product (bksp_amd) never does. ``ref()`` loads
oracle/_ref/libref_matvec.so, the reference's own src/matvec.c and src/norm.c
compiled unmodified, when it has been built in this container.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libref_matvec.so")

_o = None
_r = None
c_int, c_dbl, c_char = ctypes.c_int, ctypes.c_double, ctypes.c_char
IP = ctypes.POINTER(ctypes.c_int)
DP = ctypes.POINTER(ctypes.c_double)


def _ip(a):
    return a.ctypes.data_as(IP)


def _dp(a):
    return a.ctypes.data_as(DP)


def lib():
    global _o
    if _o is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle/liboracle.so missing: run `make -C oracle`")
        L = ctypes.CDLL(ORACLE_SO)
        for name in ("orc_ddot", "orc_dnrm2", "orc_sp_norm_f", "orc_wall_seconds", "orc_dznrm2"):
            getattr(L, name).restype = c_dbl
        _o = L
    return _o


def lib_mt():
    """The timing build (liboracle_mt.so: BLAS-1 threaded with OpenMP, SpMV serial), or
    None when it was not built. Never used as the checker."""
    p = os.path.join(HERE, "liboracle_mt.so")
    if not os.path.exists(p):
        return None
    L = ctypes.CDLL(p)
    L.orc_wall_seconds.restype = c_dbl
    return L


def cg_mt(n, rp, ci, v, b, tol, max_iter):
    """CG of the threaded timing build (same algorithm, reduction order per thread count)."""
    L = lib_mt()
    rp, ci, v = _csr(rp, ci, v)
    x = np.zeros(n)
    ret = L.orc_cg(n, _ip(rp), _ip(ci), _dp(v), _dp(np.ascontiguousarray(b, np.float64)), c_dbl(tol), max_iter,
                   _dp(x))
    return ret, x


def ref():
    """The reference's compiled kernels, or None when they were not built."""
    global _r
    if _r is None and os.path.exists(REF_SO):
        _r = ctypes.CDLL(REF_SO)
        _r.sp_norm_f.restype = c_dbl
    return _r


def _csr(rp, ci, v):
    return (np.ascontiguousarray(rp, np.int32), np.ascontiguousarray(ci, np.int32),
            np.ascontiguousarray(v, np.float64))


# ---------------------------------------------------------------- kernels

def csrgemv(trans, n, rp, ci, v, x):
    rp, ci, v = _csr(rp, ci, v)
    x = np.ascontiguousarray(x, np.float64)
    y = np.empty(n)
    lib().orc_csrgemv(c_char(trans.encode()), n, _dp(v), _ip(rp), _ip(ci), _dp(x), _dp(y))
    return y


def csrsymv(n, rp, ci, v, x):
    rp, ci, v = _csr(rp, ci, v)
    x = np.ascontiguousarray(x, np.float64)
    y = np.empty(n)
    lib().orc_csrsymv(n, _dp(v), _ip(rp), _ip(ci), _dp(x), _dp(y))
    return y


def csrmm(kind, n, s, alpha, rp, ci, v, B, beta, C):
    """mkl_dcsrmm dispatch semantics; B, C are n x s column-major (Fortran order)."""
    rp, ci, v = _csr(rp, ci, v)
    B = np.asfortranarray(B, np.float64)
    C = np.array(C, dtype=np.float64, order="F", copy=True)
    lib().orc_csrmm(c_char(kind.encode()), n, s, c_dbl(alpha), _dp(v), _ip(rp), _ip(ci),
                    _dp(B), c_dbl(beta), _dp(C))
    return C


def sp_norm_f(n, rp, ci, v, symmetric):
    rp, ci, v = _csr(rp, ci, v)
    return lib().orc_sp_norm_f(n, _ip(rp), _ip(ci), _dp(v), int(symmetric))


def dhorner(gamma, l, sigma):
    g = np.ascontiguousarray(gamma, np.float64)
    out = np.zeros(l + 1)
    lib().orc_dhorner(_dp(g), l, c_dbl(sigma), _dp(out))
    return out


def comb(n, r):
    return lib().orc_comb(n, r)


def myqr(V):
    V = np.array(V, dtype=np.float64, order="F", copy=True)
    m, s = V.shape
    R = np.zeros((s, s), order="F")
    lib().orc_myqr(m, s, _dp(V), _dp(R))
    return V, R


def ddot(x, y):
    x = np.ascontiguousarray(x, np.float64); y = np.ascontiguousarray(y, np.float64)
    return lib().orc_ddot(len(x), _dp(x), _dp(y))


def dnrm2(x):
    x = np.ascontiguousarray(x, np.float64)
    return lib().orc_dnrm2(len(x), _dp(x))


# ---------------------------------------------------------------- solvers

def _solve(fn, n, rp, ci, v, b, x0, *args):
    rp, ci, v = _csr(rp, ci, v)
    b = np.ascontiguousarray(b, np.float64)
    x = np.array(x0 if x0 is not None else np.zeros(len(b)), dtype=np.float64, copy=True)
    ret = fn(n, _ip(rp), _ip(ci), _dp(v), _dp(b), *args, _dp(x))
    return ret, x


def cg(n, rp, ci, v, b, tol, max_iter, x0=None):
    return _solve(lib().orc_cg, n, rp, ci, v, b, x0, c_dbl(tol), max_iter)


def cr(n, rp, ci, v, b, tol, max_iter, x0=None):
    return _solve(lib().orc_cr, n, rp, ci, v, b, x0, c_dbl(tol), max_iter)


def bicg(n, rp, ci, v, b, tol, max_iter, x0=None):
    return _solve(lib().orc_bicg, n, rp, ci, v, b, x0, c_dbl(tol), max_iter)


def bicgstabl(n, rp, ci, v, b, l, tol, max_iter, x0=None):
    return _solve(lib().orc_bicgstabl, n, rp, ci, v, b, x0, l, c_dbl(tol), max_iter)


def shbicgstabl(n, rp, ci, v, b, l, sigma, tol, max_iter):
    rp, ci, v = _csr(rp, ci, v)
    sig = np.ascontiguousarray(sigma, np.float64)
    x = np.zeros(n * (len(sig) + 1))
    ret = lib().orc_shbicgstabl(n, _ip(rp), _ip(ci), _dp(v), _dp(np.ascontiguousarray(b, np.float64)),
                                l, _dp(sig), len(sig), c_dbl(tol), max_iter, _dp(x))
    return ret, x.reshape(len(sig) + 1, n)


def bicgstab(n, rp, ci, v, b, tol, max_iter, x0=None):
    return _solve(lib().orc_bicgstab, n, rp, ci, v, b, x0, c_dbl(tol), max_iter)


def shbicg(n, rp, ci, v, b, sigma, tol, max_iter):
    """src/shbicg.c: returns (ret, X) with X[0] the seed and X[k+1] shift k."""
    rp, ci, v = _csr(rp, ci, v)
    sig = np.ascontiguousarray(sigma, np.float64)
    x = np.zeros(n * (len(sig) + 1))
    ret = lib().orc_shbicg(n, _ip(rp), _ip(ci), _dp(v), _dp(np.ascontiguousarray(b, np.float64)),
                           _dp(sig), len(sig), c_dbl(tol), max_iter, _dp(x))
    return ret, x.reshape(len(sig) + 1, n)


def blbicgstab(n, rp, ci, v, B, tol, max_iter, X0=None):
    rp, ci, v = _csr(rp, ci, v)
    B = np.asfortranarray(B, np.float64)
    s = B.shape[1]
    X = np.array(X0 if X0 is not None else np.zeros((n, s)), dtype=np.float64, order="F", copy=True)
    ret = lib().orc_blbicgstab(n, _ip(rp), _ip(ci), _dp(v), s, _dp(B), c_dbl(tol), max_iter, _dp(X))
    return ret, X


def blcg(n, rp, ci, v, B, tol, max_iter, X0=None, p0_mode=0):
    rp, ci, v = _csr(rp, ci, v)
    B = np.asfortranarray(B, np.float64)
    s = B.shape[1]
    X = np.array(X0 if X0 is not None else np.zeros((n, s)), dtype=np.float64, order="F", copy=True)
    ret = lib().orc_blcg(n, _ip(rp), _ip(ci), _dp(v), s, _dp(B), c_dbl(tol), max_iter, _dp(X), p0_mode)
    return ret, X


def blbicgstabl(n, rp, ci, v, B, l, tol, max_iter, X0=None):
    rp, ci, v = _csr(rp, ci, v)
    B = np.asfortranarray(B, np.float64)
    s = B.shape[1]
    X = np.array(X0 if X0 is not None else np.zeros((n, s)), dtype=np.float64, order="F", copy=True)
    tr = ctypes.c_double()
    ret = lib().orc_blbicgstabl(n, _ip(rp), _ip(ci), _dp(v), s, _dp(B), l, c_dbl(tol), max_iter,
                                _dp(X), ctypes.byref(tr))
    return ret, X, tr.value


def wall():
    return lib().orc_wall_seconds()


# ---------------------------------------------------------- complex family

def _zp(a):
    """Pointer to a complex128 array's interleaved (re, im) doubles."""
    return a.ctypes.data_as(DP)


def _zcsr(rp, ci, v):
    return (np.ascontiguousarray(rp, np.int32), np.ascontiguousarray(ci, np.int32),
            np.ascontiguousarray(v, np.complex128))


def zdotc(x, y):
    x = np.ascontiguousarray(x, np.complex128); y = np.ascontiguousarray(y, np.complex128)
    out = np.zeros(1, np.complex128)
    lib().orc_zdotc_sub(len(x), _zp(x), _zp(y), _zp(out))
    return complex(out[0])


def dznrm2(x):
    x = np.ascontiguousarray(x, np.complex128)
    return lib().orc_dznrm2(len(x), _zp(x))


def zcsrgemm0(trans, n, rp, ci, v, X):
    """C = op(A) X for X of n x s (column-major), op 'n' or 't' (conjugate transpose)."""
    rp, ci, v = _zcsr(rp, ci, v)
    X = np.asfortranarray(X, np.complex128).reshape(n, -1, order="F")
    C = np.zeros(X.shape, np.complex128, order="F")
    lib().orc_zcsrgemm0(c_char(trans.encode()), n, X.shape[1], _zp(v), _ip(rp), _ip(ci), _zp(X), _zp(C))
    return C


def zcsrgemm1(trans, alpha, n, rp, ci, v, X, beta, C):
    """C = beta C + alpha op(A) X with 0-based CSR in (shifted to 1-based here)."""
    rp, ci, v = _zcsr(rp, ci, v)
    rp1, ci1 = rp + 1, ci + 1
    X = np.asfortranarray(X, np.complex128).reshape(n, -1, order="F")
    C = np.array(C, np.complex128, order="F", copy=True).reshape(n, -1, order="F")
    al, be = np.array([alpha], np.complex128), np.array([beta], np.complex128)
    lib().orc_zcsrgemm1(c_char(trans.encode()), _zp(al), n, X.shape[1], _zp(v), _ip(rp1), _ip(ci1), _zp(X),
                        _zp(be), _zp(C))
    return C


def zhorner(gamma, l, sigma):
    g = np.ascontiguousarray(gamma, np.complex128)
    out = np.zeros(l + 1, np.complex128)
    sg = np.array([sigma], np.complex128)
    lib().orc_zhorner(_zp(g), l, _zp(sg), _zp(out))
    return out


def _zsolve(fn, n, rp, ci, v, b, x0, *args):
    rp, ci, v = _zcsr(rp, ci, v)
    b = np.ascontiguousarray(b, np.complex128)
    x = np.array(x0 if x0 is not None else np.zeros(len(b)), dtype=np.complex128, copy=True)
    ret = fn(n, _ip(rp), _ip(ci), _zp(v), _zp(b), *args, _zp(x))
    return ret, x


def zbicg(n, rp, ci, v, b, tol, max_iter, x0=None):
    return _zsolve(lib().orc_zbicg, n, rp, ci, v, b, x0, c_dbl(tol), max_iter)


def zbicgstabl(n, rp, ci, v, b, l, tol, max_iter, x0=None):
    return _zsolve(lib().orc_zbicgstabl, n, rp, ci, v, b, x0, l, c_dbl(tol), max_iter)


def zqreco(V):
    V = np.array(V, dtype=np.complex128, order="F", copy=True)
    m, s = V.shape
    R = np.zeros((s, s), np.complex128, order="F")
    lib().orc_zqreco(m, s, _zp(V), _zp(R))
    return V, R


def zblbicgstabl(n, rp, ci, v, B, l, tol, max_iter, X0=None):
    rp, ci, v = _zcsr(rp, ci, v)
    B = np.asfortranarray(B, np.complex128)
    s = B.shape[1]
    X = np.array(X0 if X0 is not None else np.zeros((n, s)), dtype=np.complex128, order="F", copy=True)
    tr = ctypes.c_double()
    ret = lib().orc_zblbicgstabl(n, _ip(rp), _ip(ci), _zp(v), s, _zp(B), l, c_dbl(tol), max_iter, _zp(X),
                                 ctypes.byref(tr))
    return ret, X, tr.value


def zshbicgstabl(n, rp, ci, v, b, l, sigma, tol, max_iter):
    rp, ci, v = _zcsr(rp, ci, v)
    sig = np.ascontiguousarray(sigma, np.complex128)
    x = np.zeros(n * (len(sig) + 1), np.complex128)
    ret = lib().orc_zshbicgstabl(n, _ip(rp), _ip(ci), _zp(v), _zp(np.ascontiguousarray(b, np.complex128)),
                                 l, _zp(sig), len(sig), c_dbl(tol), max_iter, _zp(x))
    return ret, x.reshape(len(sig) + 1, n)


# ------------------------------------------------------ reference kernels

def ref_csrgemv(trans, n, rp, ci, v, x):
    """The reference's dcsrgemm0 (src/matvec.c:11-53), for one column."""
    R = ref()
    rp, ci, v = _csr(rp, ci, v)
    x = np.ascontiguousarray(x, np.float64)
    y = np.empty(n)
    R.dcsrgemm0(c_char(trans.encode()), n, 1, _dp(v), _ip(rp), _ip(ci), int(rp[-1]), _dp(x), _dp(y))
    return y


def ref_csrsymv(n, rp, ci, v, x):
    """The reference's dcsrsymm0 (src/matvec.c:100-127) on a zeroed y, as
    mkl_cspblas_dcsrsymv prepares it (include/bksp_internal.h:39)."""
    R = ref()
    rp, ci, v = _csr(rp, ci, v)
    x = np.ascontiguousarray(x, np.float64)
    y = np.zeros(n)
    R.dcsrsymm0(c_char(b"l"), n, 1, _dp(v), _ip(rp), _ip(ci), int(rp[-1]), _dp(x), _dp(y))
    return y


def ref_csrgemm1(n, s, alpha, rp1, ci1, v, B, beta, C):
    """The reference's dcsrgemm1 'n' (src/matvec.c:130-172) on 1-based arrays."""
    R = ref()
    rp1 = np.ascontiguousarray(rp1, np.int32); ci1 = np.ascontiguousarray(ci1, np.int32)
    v = np.ascontiguousarray(v, np.float64)
    B = np.asfortranarray(B, np.float64)
    C = np.array(C, dtype=np.float64, order="F", copy=True)
    R.dcsrgemm1(c_char(b"n"), c_dbl(alpha), n, s, _dp(v), _ip(rp1), _ip(ci1), int(rp1[-1] - 1),
                _dp(B), c_dbl(beta), _dp(C))
    return C


def ref_zcsrgemm0(trans, n, rp, ci, v, X):
    """The reference's zcsrgemm0 (src/matvec.c:55-98) on 0-based CSR, X n x s column-major."""
    R = ref()
    rp, ci, v = _zcsr(rp, ci, v)
    X = np.asfortranarray(X, np.complex128).reshape(n, -1, order="F")
    C = np.zeros(X.shape, np.complex128, order="F")
    R.zcsrgemm0(c_char(trans.encode()), n, X.shape[1], _zp(v), _ip(rp), _ip(ci), int(rp[-1]), _zp(X), _zp(C))
    return C


def ref_zcsrgemm1(trans, alpha, n, rp, ci, v, X, beta, C):
    """The reference's zcsrgemm1 (src/matvec.c:174-217); 0-based CSR in, 1-based passed."""
    R = ref()
    rp, ci, v = _zcsr(rp, ci, v)
    rp1, ci1 = rp + 1, ci + 1
    X = np.asfortranarray(X, np.complex128).reshape(n, -1, order="F")
    C = np.array(C, np.complex128, order="F", copy=True).reshape(n, -1, order="F")
    al, be = np.array([alpha], np.complex128), np.array([beta], np.complex128)
    R.zcsrgemm1(c_char(trans.encode()), _zp(al), n, X.shape[1], _zp(v), _ip(rp1), _ip(ci1), int(rp[-1]), _zp(X),
                _zp(be), _zp(C))
    return C


def ref_csrsymm1(n, s, alpha, rp1, ci1, v, B, C):
    """The reference's dcsrsymm1 'n' (src/matvec.c:219-246) accumulating into C."""
    R = ref()
    rp1 = np.ascontiguousarray(rp1, np.int32); ci1 = np.ascontiguousarray(ci1, np.int32)
    v = np.ascontiguousarray(v, np.float64)
    B = np.asfortranarray(B, np.float64)
    C = np.array(C, dtype=np.float64, order="F", copy=True)
    R.dcsrsymm1(c_char(b"n"), c_dbl(alpha), n, s, _dp(v), _ip(rp1), _ip(ci1), int(rp1[-1] - 1),
                _dp(B), c_dbl(0.0), _dp(C))
    return C


def ref_sp_norm_f(n, rp, ci, v, symmetric):
    """The reference's sp_norm_f (src/norm.c:37-47)."""
    R = ref()

    class _Mat(ctypes.Structure):
        _fields_ = [("row", IP), ("col", IP), ("nnz", IP), ("row_size", IP), ("col_size", IP),
                    ("value", DP), ("format", ctypes.c_char_p), ("type", ctypes.c_char_p)]
    rp, ci, v = _csr(rp, ci, v)
    nn, nz = ctypes.c_int(n), ctypes.c_int(int(rp[-1]))
    m = _Mat(_ip(rp), _ip(ci), ctypes.pointer(nz), ctypes.pointer(nn), ctypes.pointer(nn), _dp(v),
             b"csr", b"symmetric" if symmetric else b"unsymmetric")
    return R.sp_norm_f(ctypes.byref(m))


# ------------------------------------------------- learned FSAI (orc_fsai.c)

FP = ctypes.POINTER(ctypes.c_float)


def _fp(a):
    return a.ctypes.data_as(FP)


def fsai_nparam():
    return lib().orc_fsai_nparam()


def fsai_rows(n, rp, ci, v):
    """sqrt(a_ii), normalised coupling sums, lower-entry counts, diagonal excess of
    the expanded symmetric matrix."""
    rp, ci, v = _csr(rp, ci, v)
    out = [np.empty(n) for _ in range(4)]
    lib().orc_fsai_rows(n, _ip(rp), _ip(ci), _dp(v), *[_dp(o) for o in out])
    return tuple(out)


def fsai_features(n, rpl, cil, vl, rows):
    rpl, cil, vl = _csr(rpl, cil, vl)
    sqd, s, nlow, ex = rows
    F = np.empty((int(rpl[-1]), 8), np.float32)
    f = np.empty(8, np.float32)
    for i in range(n):
        for k in range(rpl[i], rpl[i + 1]):
            lib().orc_fsai_edge_features(n, i, int(cil[k]), c_dbl(vl[k]), _dp(sqd), _dp(s), _dp(nlow),
                                         _dp(ex), _fp(f))
            F[k] = f
    return F


def fsai_forward(n, rpl, cil, vl, rows, params, bf16):
    rpl, cil, vl = _csr(rpl, cil, vl)
    sqd, s, nlow, ex = rows
    p = np.ascontiguousarray(params, np.float32)
    h = np.empty(int(rpl[-1]))
    lib().orc_fsai_forward(n, _ip(rpl), _ip(cil), _dp(vl), _dp(sqd), _dp(s), _dp(nlow), _dp(ex), _fp(p),
                           int(bf16), _dp(h))
    return h


def fsai_gvals(n, rpl, cil, sqd, h):
    rpl, cil = np.ascontiguousarray(rpl, np.int32), np.ascontiguousarray(cil, np.int32)
    g = np.empty(int(rpl[-1]))
    lib().orc_fsai_gvals(n, _ip(rpl), _ip(cil), _dp(np.ascontiguousarray(sqd)), _dp(np.ascontiguousarray(h)),
                         _dp(g))
    return g


def fsai_loss_dg(n, rp, ci, v, rpl, cil, g, Z):
    rp, ci, v = _csr(rp, ci, v)
    rpl, cil = np.ascontiguousarray(rpl, np.int32), np.ascontiguousarray(cil, np.int32)
    Z = np.ascontiguousarray(Z, np.float64)
    g = np.ascontiguousarray(g, np.float64)
    dg = np.empty(int(rpl[-1]))
    lib().orc_fsai_loss_dg.restype = c_dbl
    L = lib().orc_fsai_loss_dg(n, _ip(rp), _ip(ci), _dp(v), _ip(rpl), _ip(cil), _dp(g), Z.shape[0], _dp(Z),
                               _dp(dg))
    return L, dg


def fsai_backward(n, rpl, cil, vl, rows, params, bf16, h, dg):
    rpl, cil, vl = _csr(rpl, cil, vl)
    sqd, s, nlow, ex = rows
    p = np.ascontiguousarray(params, np.float32)
    grad = np.empty(fsai_nparam())
    lib().orc_fsai_backward(n, _ip(rpl), _ip(cil), _dp(vl), _dp(sqd), _dp(s), _dp(nlow), _dp(ex), _fp(p),
                            int(bf16), _dp(np.ascontiguousarray(h)), _dp(np.ascontiguousarray(dg)),
                            _dp(grad))
    return grad


def fsai_loss_grad(n, rp, ci, v, rpl, cil, vl, params, Z, bf16=False):
    """Loss and parameter gradient of the model on one field (A expanded rp/ci/v,
    its lower triangle rpl/cil/vl) with probes Z (P x n)."""
    rows = fsai_rows(n, rp, ci, v)
    h = fsai_forward(n, rpl, cil, vl, rows, params, bf16)
    g = fsai_gvals(n, rpl, cil, rows[0], h)
    L, dg = fsai_loss_dg(n, rp, ci, v, rpl, cil, g, Z)
    return L, fsai_backward(n, rpl, cil, vl, rows, params, bf16, h, dg), g
