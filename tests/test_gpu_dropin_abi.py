"""The remaining drop-in symbols of the reference library on the GPU:
read_matrix_market, sp_norm_f, myqr, the four CSR kernels and the dmat helpers,
checked against the oracle (and the reference's own matvec.c/norm.c where built)."""
import ctypes
import os

import numpy as np
import pytest

from bksp_amd import gen, load
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


def _myqr(V):
    m, s = V.shape
    Q = np.asfortranarray(V.copy())
    R = np.full((s, s), np.nan, order="F")
    rc = load().myqr(m, s, dp(Q), dp(R))
    return rc, Q, R


@pytest.mark.parametrize("m,s", [(1000, 1), (1000, 5), (1000, 8), (1000, 16), (300, 40), (50, 50), (70000, 8)])
def test_myqr_lapack_convention(m, s):
    """src/myqr.c:4-39: dgeqrf + dorgqr. The device Householder QR returns LAPACK's Q
    and R, signs included (numpy's qr is LAPACK dgeqrf/dorgqr), to 1e-12; the oracle's
    restatement agrees too. Square and wider-than-16 panels included."""
    rng = np.random.default_rng(52 + s)
    V = np.asfortranarray(rng.standard_normal((m, s)))
    rc, Q, R = _myqr(V)
    assert rc == 0
    Qn, Rn = np.linalg.qr(V)
    assert np.abs(R - Rn).max() <= 1e-12 * np.abs(Rn).max()
    assert np.abs(Q - Qn).max() <= 1e-12
    assert np.array_equal(np.tril(R, -1), np.zeros((s, s)))
    Qo, Ro = po.myqr(V)
    assert np.abs(R - Ro).max() <= 1e-12 * np.abs(Ro).max() and np.abs(Q - Qo).max() <= 1e-12


def test_myqr_ill_conditioned_and_rank_deficient():
    """LAPACK's Householder QR returns 0 for any panel; so does the device one. cond(V)
    = 1e10: Q orthonormal, QR = V, R equal to LAPACK's (diagonal signs included). A
    repeated column: R's diagonal entry for it is at rounding level, Q stays orthonormal
    and QR = V (a Cholesky-based QR breaks down on both)."""
    rng = np.random.default_rng(7)
    m, s = 2000, 8
    U, _ = np.linalg.qr(rng.standard_normal((m, s)))
    W, _ = np.linalg.qr(rng.standard_normal((s, s)))
    V = np.asfortranarray(U @ np.diag(np.logspace(0, -10, s)) @ W.T)
    rc, Q, R = _myqr(V)
    assert rc == 0
    Qn, Rn = np.linalg.qr(V)
    assert np.abs(Q.T @ Q - np.eye(s)).max() < 1e-13
    assert np.abs(Q @ R - V).max() <= 1e-14 * np.abs(V).max() * 10
    assert np.abs(R - Rn).max() <= 1e-12 * np.abs(Rn).max()
    assert np.array_equal(np.sign(np.diag(R)), np.sign(np.diag(Rn)))
    V = np.asfortranarray(rng.standard_normal((m, s)))
    V[:, 5] = V[:, 2]
    rc, Q, R = _myqr(V)
    assert rc == 0
    assert abs(R[5, 5]) < 1e-12 * np.abs(R).max()
    assert np.abs(Q.T @ Q - np.eye(s)).max() < 1e-13
    assert np.abs(Q @ R - V).max() < 1e-12
    Qn, Rn = np.linalg.qr(V)
    assert np.abs(R[:5] - Rn[:5]).max() <= 1e-12 * np.abs(Rn).max()


def test_myqr_argument_codes():
    """dorgqr's info for n > m and m < 0 (myqr returns it, src/myqr.c:30-38)."""
    V = np.asfortranarray(np.ones((3, 4)))
    R = np.zeros((4, 4), order="F")
    assert load().myqr(3, 4, dp(V), dp(R)) == -2
    assert load().myqr(-1, 2, dp(V), dp(R)) == -1
    assert load().myqr(3, 0, dp(V), dp(R)) == 0


def test_dhorner_comb_and_dmat():
    gam = np.array([0.3, -0.2, 0.05, 0.1])
    out = np.zeros(5)
    load().dhorner(dp(gam), 4, 0.37, dp(out))
    assert np.allclose(out, po.dhorner(gam, 4, 0.37), rtol=0, atol=0)
    assert load().comb(7, 3) == 35 and load().comb(3, 5) == 0
    M = load().dmat_create(4, 3).contents
    assert M.row_size == 4 and M.col_size == 3
    assert all(M.value[i] == 0.0 for i in range(12))


def test_dropin_device_copy_cache(monkeypatch):
    """bksp_dcg keeps the device copy of A per thread, keyed by the dspmat's arrays and
    a hash of their contents: repeated calls give the same x as an uncached call, and a
    caller that rewrites values in place gets a solve of the new matrix."""
    from bksp_amd import bksp
    n, rp, ci, v = gen.poisson3d(24, lower=True)
    v = v.copy()
    b = gen.normal(3, n)
    A = bksp.dspmat(n, rp, ci, v, symmetric=True)
    xs = []
    for _ in range(3):
        x = np.zeros(n)
        assert bksp.bksp_dcg(A, b, 1e-10, 2000, x) >= 0
        xs.append(x)
    monkeypatch.setenv("BKSP_CACHE", "0")
    x0 = np.zeros(n)
    assert bksp.bksp_dcg(A, b, 1e-10, 2000, x0) >= 0
    monkeypatch.delenv("BKSP_CACHE")
    assert np.array_equal(xs[0], xs[1]) and np.array_equal(xs[1], xs[2]) and np.array_equal(xs[0], x0)
    # rewrite the diagonal in place (the same arrays, new contents): the cache must miss
    diag = np.array([k for i in range(n) for k in range(rp[i], rp[i + 1]) if ci[k] == i])
    A.val[diag] *= 2.0
    x = np.zeros(n)
    ret = bksp.bksp_dcg(A, b, 1e-10, 2000, x)
    cret, cx = po.cg(n, rp, ci, A.val, b, 1e-10, 2000)
    assert ret >= 0 and abs(ret - cret) <= max(1, int(0.05 * cret)), (ret, cret)
    assert np.linalg.norm(x - cx) / np.linalg.norm(cx) < 1e-8
