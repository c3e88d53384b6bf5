"""Direct pins of the oracle's LU routines and of its block solvers for s > 1.

orc_dgetrf / orc_dgetrs / orc_dgesv / orc_dgetri restate LAPACK's dgetrf, dgetrs,
dgesv and dgetri (the reference calls them through LAPACKE in blbicgstab.c:69,94,
blbicgstabl.c and blcg.c:85-93). They are checked here on their own, for n > 1:
P A = L U with unit-lower L and LAPACK's 1-based row interchanges, solves against
numpy's, the inverse against the identity, and the singular-pivot info code.

The block solvers are pinned for s > 1 by finite termination, which only a correct
block recurrence has: with n = s * l (the block Krylov space of degree l fills R^n)
one outer step of block BiCGSTAB(l) is exact, block BiCGSTAB (blbicgstab.c) is exact
after one step for n = s, and block CG (P0 = V) after k steps for n = s * k. Each
case has a negative control one row larger, where the same step count leaves a
residual many orders above rounding: a dropped term, a transposed s x s operand or a
wrong sign in the block coefficients breaks the exact case."""
import ctypes

import numpy as np
import pytest

from oracle import pyoracle as po

c_int, c_dbl = ctypes.c_int, ctypes.c_double


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(c_int))


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(c_dbl))


def getrf(A):
    n = A.shape[0]
    LU = np.asfortranarray(A, dtype=np.float64).copy(order="F")
    piv = np.zeros(n, np.int32)
    info = po.lib().orc_dgetrf(n, _dp(LU), n, _ip(piv))
    return info, LU, piv


def perm_from_ipiv(piv):
    """Row order that LAPACK's sequential interchanges (1-based ipiv) apply."""
    p = np.arange(len(piv))
    for k, q in enumerate(piv - 1):
        p[[k, q]] = p[[q, k]]
    return p


@pytest.mark.parametrize("n", [1, 2, 5, 16])
def test_dgetrf_is_pa_equals_lu(n):
    rng = np.random.default_rng(100 + n)
    A = rng.standard_normal((n, n))
    info, LU, piv = getrf(A)
    assert info == 0
    assert np.all((piv >= np.arange(1, n + 1)) & (piv <= n))     # 1-based, never above the diagonal
    L = np.tril(LU, -1) + np.eye(n)
    U = np.triu(LU)
    assert np.all(np.abs(np.tril(LU, -1)) <= 1.0 + 1e-15)         # partial pivoting bounds |l_ij| <= 1
    assert np.allclose(A[perm_from_ipiv(piv)], L @ U, rtol=0, atol=1e-13 * np.abs(A).max() * n)


def test_dgetrf_pivots_a_zero_leading_entry():
    A = np.array([[0.0, 2.0, 1.0], [3.0, 1.0, 0.0], [1.0, 1.0, 4.0]])
    info, LU, piv = getrf(A)
    assert info == 0 and piv[0] == 2                               # row 2 (1-based) holds the largest |a_i1|
    L = np.tril(LU, -1) + np.eye(3)
    assert np.allclose(A[perm_from_ipiv(piv)], L @ np.triu(LU), atol=1e-15)


def test_dgetrf_singular_reports_the_zero_pivot():
    A = np.array([[1.0, 2.0, 3.0], [2.0, 4.0, 6.0], [1.0, 0.0, 1.0]])   # row 2 = 2 * row 1
    info, _, _ = getrf(A)
    assert info > 0                                                # LAPACK: U(info, info) is exactly zero


@pytest.mark.parametrize("n,nrhs", [(2, 1), (6, 3), (16, 16)])
def test_dgetrs_and_dgesv_solve(n, nrhs):
    rng = np.random.default_rng(7 * n + nrhs)
    A = rng.standard_normal((n, n)) + n * np.eye(n)
    B = rng.standard_normal((n, nrhs))
    ref = np.linalg.solve(A, B)
    info, LU, piv = getrf(A)
    X = np.asfortranarray(B).copy(order="F")
    assert po.lib().orc_dgetrs(n, nrhs, _dp(LU), n, _ip(piv), _dp(X), n) == 0
    assert np.allclose(X, ref, rtol=1e-12, atol=1e-12)
    A2 = np.asfortranarray(A).copy(order="F")
    X2 = np.asfortranarray(B).copy(order="F")
    piv2 = np.zeros(n, np.int32)
    assert po.lib().orc_dgesv(n, nrhs, _dp(A2), n, _ip(piv2), _dp(X2), n) == 0
    assert np.array_equal(X2, X)                                   # dgesv = dgetrf + dgetrs
    assert np.array_equal(piv2, piv)


@pytest.mark.parametrize("n", [3, 8])
def test_dgetri_inverse(n):
    rng = np.random.default_rng(50 + n)
    A = rng.standard_normal((n, n)) + n * np.eye(n)
    info, LU, piv = getrf(A)
    assert po.lib().orc_dgetri(n, _dp(LU), n, _ip(piv)) == 0
    assert np.allclose(A @ LU, np.eye(n), atol=1e-12)


# ---- block solvers, s > 1: finite termination ----

def dense_csr(A):
    n = A.shape[0]
    rp = np.arange(0, n * n + 1, n, dtype=np.int32)
    ci = np.tile(np.arange(n, dtype=np.int32), n)
    return n, rp, ci, np.ascontiguousarray(A).ravel().copy()


def lower_csr(A):
    n = A.shape[0]
    rows, cols = np.tril_indices(n)
    rp = np.concatenate([[0], np.cumsum(np.arange(1, n + 1))]).astype(np.int32)
    return n, rp, cols.astype(np.int32), A[rows, cols].copy()


def general(n, seed):
    rng = np.random.default_rng(seed)
    return 3.0 * np.eye(n) + 0.3 / np.sqrt(n) * rng.standard_normal((n, n))


def spd(n, seed):
    rng = np.random.default_rng(seed)
    Q = rng.standard_normal((n, n))
    return Q @ Q.T / n + np.eye(n)


def relres(A, B, X):
    return np.linalg.norm(B - A @ X) / np.linalg.norm(B)


@pytest.mark.parametrize("s,l", [(2, 1), (2, 2), (4, 2), (3, 3), (8, 2)])
def test_blbicgstabl_one_outer_step_is_exact_when_n_is_s_l(s, l):
    for n, exact in ((s * l, True), (s * l + 1, False)):
        A = general(n, 10 * s + l)
        B = np.asfortranarray(np.random.default_rng(n).standard_normal((n, s)))
        _, X, _ = po.blbicgstabl(*dense_csr(A), B, l, 1e-300, 1)
        r = relres(A, B, X)
        assert (r < 1e-13) if exact else (r > 1e-9), (s, l, n, r)


@pytest.mark.parametrize("s", [2, 3, 4, 8])
def test_blbicgstab_one_step_is_exact_when_n_is_s(s):
    for n, exact in ((s, True), (s + 1, False)):
        A = general(n, 20 + s)
        B = np.asfortranarray(np.random.default_rng(n).standard_normal((n, s)))
        _, X = po.blbicgstab(*dense_csr(A), B, 1e-300, 1)
        r = relres(A, B, X)
        assert (r < 1e-12) if exact else (r > 1e-6), (s, n, r)


@pytest.mark.parametrize("s,k", [(2, 3), (4, 2), (8, 1), (8, 3)])
def test_blcg_k_steps_are_exact_when_n_is_s_k(s, k):
    """Corrected start P0 = V (p0_mode 1); the reference's P0 = 2V start is pinned by
    its half first step in test_oracle_solvers."""
    for n, exact in ((s * k, True), (s * k + 1, False)):
        A = spd(n, 30 + s + k)
        B = np.asfortranarray(np.random.default_rng(n).standard_normal((n, s)))
        _, X = po.blcg(*lower_csr(A), B, 1e-300, k, p0_mode=1)
        r = relres(A, B, X)
        assert (r < 1e-13) if exact else (r > 1e-5), (s, k, n, r)
