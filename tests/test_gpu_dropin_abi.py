"""The remaining drop-in symbols of the reference library on the GPU:
read_matrix_market, sp_norm_f, myqr, the four CSR kernels and the dmat helpers,
checked against the oracle (and the reference's own matvec.c/norm.c where built)."""
import ctypes
import os

import numpy as np
import pytest

from bksp_amd import load
from bksp_amd._lib import DMat
from conftest import GOLDEN, read_mtx
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu
IP = ctypes.POINTER(ctypes.c_int)
DP = ctypes.POINTER(ctypes.c_double)


def dp(a):
    return a.ctypes.data_as(DP)


def ip(a):
    return a.ctypes.data_as(IP)


@pytest.mark.parametrize("name", ["bfwb398.mtx", "pde2961.mtx"])
def test_read_matrix_market_and_norm(name):
    path = os.path.join(GOLDEN, "testmat", name)
    A = load().read_matrix_market(path.encode()).contents
    n = A.row_size[0]
    nnz = A.nnz[0]
    rp = np.ctypeslib.as_array(A.row, shape=(n + 1,))
    ci = np.ctypeslib.as_array(A.col, shape=(nnz,))
    v = np.ctypeslib.as_array(A.value, shape=(nnz,))
    n2, rp2, ci2, v2, sym = read_mtx(path)
    assert n == n2 and np.array_equal(rp, rp2) and np.array_equal(ci, ci2) and np.array_equal(v, v2)
    assert A.type.decode() == ("symmetric" if sym else "unsymmetric") and A.format.decode() == "csr"
    got = load().sp_norm_f(ctypes.byref(A))
    want = po.sp_norm_f(n, rp, ci, v, sym)
    assert abs(got - want) <= 1e-14 * want
    if po.ref() is not None:
        assert abs(got - po.ref_sp_norm_f(n, rp, ci, v, sym)) <= 1e-14 * want


def test_csr_kernel_symbols_bitwise(pde2961, bfwb398):
    n, rp, ci, v, _ = pde2961
    rng = np.random.default_rng(51)
    B = np.asfortranarray(rng.standard_normal((n, 2)))
    for t in "nt":
        C = np.zeros((n, 2), order="F")
        assert load().dcsrgemm0(t.encode(), n, 2, dp(v), ip(rp), ip(ci), int(rp[-1]), dp(B), dp(C)) == 0
        for j in range(2):
            assert np.array_equal(C[:, j], po.csrgemv(t, n, rp, ci, v, B[:, j]))
    C0 = np.asfortranarray(rng.standard_normal((n, 2)))
    C = C0.copy(order="F")
    rp1, ci1 = (rp + 1).astype(np.int32), (ci + 1).astype(np.int32)
    assert load().dcsrgemm1(b"n", -0.5, n, 2, dp(v), ip(rp1), ip(ci1), int(rp[-1]), dp(B), 2.0, dp(C)) == 0
    assert np.array_equal(C, po.csrmm("G", n, 2, -0.5, rp, ci, v, B, 2.0, C0))
    n, rp, ci, v, _ = bfwb398
    B = np.asfortranarray(rng.standard_normal((n, 3)))
    C = np.zeros((n, 3), order="F")
    assert load().dcsrsymm0(b"l", n, 3, dp(v), ip(rp), ip(ci), int(rp[-1]), dp(B), dp(C)) == 0
    for j in range(3):
        assert np.array_equal(C[:, j], po.csrsymv(n, rp, ci, v, B[:, j]))
    C = np.zeros((n, 3), order="F")
    rp1, ci1 = (rp + 1).astype(np.int32), (ci + 1).astype(np.int32)
    assert load().dcsrsymm1(b"n", 1.0, n, 3, dp(v), ip(rp1), ip(ci1), int(rp[-1]), dp(B), 0.0, dp(C)) == 0
    assert np.array_equal(C, po.csrmm("S", n, 3, 1.0, rp, ci, v, B, 0.0, np.zeros((n, 3))))
    assert load().dcsrgemm0(b"x", n, 1, dp(v), ip(rp), ip(ci), 0, dp(B), dp(C)) == -1


@pytest.mark.parametrize("s", [1, 5, 8, 16])
def test_myqr(s):
    rng = np.random.default_rng(52 + s)
    V = np.asfortranarray(rng.standard_normal((1000, s)))
    Q = V.copy(order="F")
    R = np.zeros((s, s), order="F")
    assert load().myqr(1000, s, dp(Q), dp(R)) == 0
    assert np.allclose(Q.T @ Q, np.eye(s), atol=1e-13)
    assert np.allclose(Q @ R, V, atol=1e-12)
    assert np.allclose(np.tril(R, -1), 0) and np.all(np.diag(R) > 0)
    _, Ro = po.myqr(V)      # LAPACK signs: |R| agrees
    assert np.allclose(np.abs(R), np.abs(Ro), atol=1e-11)


def test_dhorner_comb_and_dmat():
    gam = np.array([0.3, -0.2, 0.05, 0.1])
    out = np.zeros(5)
    load().dhorner(dp(gam), 4, 0.37, dp(out))
    assert np.allclose(out, po.dhorner(gam, 4, 0.37), rtol=0, atol=0)
    assert load().comb(7, 3) == 35 and load().comb(3, 5) == 0
    M = load().dmat_create(4, 3).contents
    assert M.row_size == 4 and M.col_size == 3
    assert all(M.value[i] == 0.0 for i in range(12))
