"""Front-end with the MATLAB gateways' contract (reference matlab/bicg.c,
bicgstabl.c, blbicgstabl.c, shbicgstabl.c, util.c), for Python callers.

Each function takes a square ``scipy.sparse`` matrix and dense right-hand sides,
checks its arguments as the gateway does (same identifiers and messages, raised
as ``GatewayError``), converts A to CSR with ascending columns in every row (the
gateway transposes MATLAB's CSC, util.c:6-34), picks the real solver when A is
real and the complex one when A is complex (``mxGetPi(A) == NULL``; for a real A
the imaginary part of b is ignored, as the gateway reads only ``mxGetPr``), starts
from x = 0, prints the gateway's convergence message and returns x (n, or n x s for
the block solver, or n x (number of shifts + 1) for the shifted solver, seed
first). The solve itself runs on the GPU through libbksp_amd.so.
"""
import numpy as np

from . import bksp


class GatewayError(ValueError):
    """A gateway argument error: ``ident`` is the gateway's mexErrMsgIdAndTxt id."""

    def __init__(self, ident, msg):
        super().__init__(f"{ident}: {msg}")
        self.ident, self.msg = ident, msg


def _scalar(v):
    return np.ndim(v) == 0 or np.size(v) == 1


def _check(name, A, b, kind, nargs, nwant, words):
    import scipy.sparse as sp
    if nargs != nwant:
        raise GatewayError(f"BKSP:{name}", f"{words} inputs required.")
    if not sp.issparse(A) or A.shape[0] != A.shape[1]:                 # util.c:85-89
        raise GatewayError("BKSP:invalidInputSparisty", "Input argument must be a sparse square matrix.")
    b = np.asarray(b)
    nvec = 1 if b.ndim == 1 else b.shape[1]
    if kind == "single" and nvec != 1:                                    # util.c:91-96
        raise GatewayError("BKSP:invalidInputVector", "Input right hand side must be a vector.")
    if b.shape[0] != A.shape[0]:                                          # util.c:100-104
        raise GatewayError("BKSP:invalidInputVector",
                           "Input right hand side must have the same size of the coefficient matrix.")


def _check_tol_maxit(tol, maxit):
    if not _scalar(tol):                                                  # util.c:107-113
        raise GatewayError("BKSP:blbicgstabl", "Input tolerance must be a scalar.")
    if not _scalar(maxit):                                                # util.c:115-121
        raise GatewayError("BKSP:blbicgstabl", "Input max iteration must be an integer.")


def _check_ell(name, ell):
    if not _scalar(ell):
        raise GatewayError(f"BKSP:{name}", "Input parameter ell must be an integer.")


def _csr(A):
    C = A.tocsr().sorted_indices()
    return C.shape[0], C.indptr.astype(np.int32), C.indices.astype(np.int32), C.data


def _is_complex(A):
    return np.iscomplexobj(A.data if hasattr(A, "data") else A)


def _report(what, code):
    if code < 0:
        print(f"The {what} Method did not converged.")
    else:
        print(f"The {what} Method converged at iteration {code}.")


def bicg(*args):
    """x = bicg(A, b, tol, maxit)  (matlab/bicg.c)."""
    A, b = (args + (None, None))[:2]
    _check("bicg", A, b, "single", len(args), 4, "Four")
    tol, maxit = args[2], args[3]
    _check_tol_maxit(tol, maxit)
    n, rp, ci, v = _csr(A)
    if _is_complex(A):
        x = np.zeros(n, np.complex128)
        code = bksp.zbicg(bksp.zspmat(n, rp, ci, v), np.ascontiguousarray(np.ravel(b), np.complex128),
                          float(np.ravel(tol)[0]), int(np.ravel(maxit)[0]), x)
    else:
        x = np.zeros(n)
        code = bksp.dbicg(bksp.dspmat(n, rp, ci, v), np.ascontiguousarray(np.real(np.ravel(b)), np.float64),
                          float(np.ravel(tol)[0]), int(np.ravel(maxit)[0]), x)
    _report("BiCG", code)
    return x


def bicgstabl(*args):
    """x = bicgstabl(A, b, ell, tol, maxit)  (matlab/bicgstabl.c)."""
    A, b = (args + (None, None))[:2]
    _check("bicgstabl", A, b, "single", len(args), 5, "Five")
    ell, tol, maxit = args[2], args[3], args[4]
    _check_ell("bicgstabl", ell)
    _check_tol_maxit(tol, maxit)
    n, rp, ci, v = _csr(A)
    l, t, mi = int(np.ravel(ell)[0]), float(np.ravel(tol)[0]), int(np.ravel(maxit)[0])
    if _is_complex(A):
        x = np.zeros(n, np.complex128)
        code = bksp.zbicgstabl(bksp.zspmat(n, rp, ci, v), np.ascontiguousarray(np.ravel(b), np.complex128),
                               l, t, mi, x)
    else:
        x = np.zeros(n)
        code = bksp.dbicgstabl(bksp.dspmat(n, rp, ci, v), np.ascontiguousarray(np.real(np.ravel(b)), np.float64),
                               l, t, mi, x)
    _report("BiCGSTAB(l)", code)
    return x


def blbicgstabl(*args):
    """X = blbicgstabl(A, B, ell, tol, maxit)  (matlab/blbicgstabl.c)."""
    A, B = (args + (None, None))[:2]
    _check("blbicgstabl", A, B, "block", len(args), 5, "Five")
    ell, tol, maxit = args[2], args[3], args[4]
    _check_ell("blbicgstabl", ell)
    _check_tol_maxit(tol, maxit)
    n, rp, ci, v = _csr(A)
    B = np.asarray(B)
    B2 = B.reshape(n, -1)
    l, t, mi = int(np.ravel(ell)[0]), float(np.ravel(tol)[0]), int(np.ravel(maxit)[0])
    if _is_complex(A):
        Bm, Xm = bksp.ZMatrix(B2.astype(np.complex128)), bksp.ZMatrix(np.zeros(B2.shape, np.complex128))
        code = bksp.zblbicgstabl(bksp.zspmat(n, rp, ci, v), Bm, l, t, mi, Xm)
    else:
        Bm, Xm = bksp.Mat(np.real(B2)), bksp.Mat(np.zeros(B2.shape))
        code = bksp.dblbicgstabl(bksp.dspmat(n, rp, ci, v), Bm, l, t, mi, Xm)
    _report("Block BiCGSTAB(l)", code)
    return Xm.a


def shbicgstabl(*args):
    """X = shbicgstabl(A, b, ell, sigma, tol, maxit)  (matlab/shbicgstabl.c); column k+1
    of X solves (A + sigma_k I) x = b, column 0 is the seed."""
    A, b = (args + (None, None))[:2]
    _check("shbicgstabl", A, b, "single", len(args), 6, "Six")
    ell, sigma, tol, maxit = args[2], args[3], args[4], args[5]
    _check_ell("shbicgstabl", ell)
    _check_tol_maxit(tol, maxit)
    n, rp, ci, v = _csr(A)
    sig = np.ravel(sigma)
    l, t, mi = int(np.ravel(ell)[0]), float(np.ravel(tol)[0]), int(np.ravel(maxit)[0])
    if _is_complex(A):
        x = np.zeros(n * (len(sig) + 1), np.complex128)
        code = bksp.zshbicgstabl(bksp.zspmat(n, rp, ci, v), np.ascontiguousarray(np.ravel(b), np.complex128), l,
                                 sig.astype(np.complex128), t, mi, x)
    else:
        x = np.zeros(n * (len(sig) + 1))
        code = bksp.dshbicgstabl(bksp.dspmat(n, rp, ci, v), np.ascontiguousarray(np.real(np.ravel(b)), np.float64),
                                 l, np.real(sig).astype(np.float64), t, mi, x)
    _report("Shifted BiCGSTAB(l)", code)
    return x.reshape(len(sig) + 1, n).T
