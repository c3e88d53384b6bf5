"""Pins the oracle's kernels and helpers (CPU only).

SpMV/SpMM: bitwise against the reference's own src/matvec.c (oracle/_ref, when
built in this container), against scipy.sparse and dense products, and against
the closed-form eigenpairs of the Dirichlet Laplacian. Helpers: Horner
re-expansion against direct polynomial evaluation, comb against math.comb, QR
against its invariants and LAPACK's signs (numpy), sp_norm_f against the
reference and a dense Frobenius norm.
"""
import math

import numpy as np
import pytest

from conftest import expand_lower, to_scipy
from oracle import pyoracle as po

needs_ref = pytest.mark.skipif(po.ref() is None, reason="oracle/_ref not built (needs /root/reference)")


def rand_csr(n, per_row, seed, square=True):
    rng = np.random.default_rng(seed)
    rp = [0]
    ci, v = [], []
    for i in range(n):
        k = rng.integers(1, per_row + 1)
        cols = rng.choice(n, size=k, replace=False)
        rng.shuffle(cols)                      # unsorted rows exercise the stored order
        ci += list(cols)
        v += list(rng.standard_normal(k))
        rp.append(len(ci))
    return n, np.array(rp, np.int32), np.array(ci, np.int32), np.array(v)


def rand_lower(n, per_row, seed):
    rng = np.random.default_rng(seed)
    rp, ci, v = [0], [], []
    for i in range(n):
        k = min(i, rng.integers(0, per_row))
        cols = sorted(rng.choice(i, size=k, replace=False)) if k else []
        ci += list(cols) + [i]
        v += list(rng.standard_normal(k)) + [float(per_row * 2)]
        rp.append(len(ci))
    return n, np.array(rp, np.int32), np.array(ci, np.int32), np.array(v)


@needs_ref
def test_gemv_bitwise_vs_reference(pde2961):
    n, rp, ci, v, _ = pde2961
    x = np.random.default_rng(1).standard_normal(n)
    for t in "nt":
        assert np.array_equal(po.csrgemv(t, n, rp, ci, v, x), po.ref_csrgemv(t, n, rp, ci, v, x))
    n, rp, ci, v = rand_csr(500, 9, 2)
    x = np.random.default_rng(3).standard_normal(n)
    for t in "nt":
        assert np.array_equal(po.csrgemv(t, n, rp, ci, v, x), po.ref_csrgemv(t, n, rp, ci, v, x))


@needs_ref
def test_symv_bitwise_vs_reference(bfwb398):
    n, rp, ci, v, sym = bfwb398
    assert sym
    x = np.random.default_rng(4).standard_normal(n)
    assert np.array_equal(po.csrsymv(n, rp, ci, v, x), po.ref_csrsymv(n, rp, ci, v, x))


@needs_ref
def test_csrmm_bitwise_vs_reference(pde2961, bfwb398):
    n, rp, ci, v, _ = pde2961
    s = 5
    rng = np.random.default_rng(5)
    B = np.asfortranarray(rng.standard_normal((n, s)))
    C = np.asfortranarray(rng.standard_normal((n, s)))
    got = po.csrmm("G", n, s, -1.0, rp, ci, v, B, 0.5, C)
    want = po.ref_csrgemm1(n, s, -1.0, rp + 1, ci + 1, v, B, 0.5, C)
    assert np.array_equal(got, want)
    n, rp, ci, v, _ = bfwb398
    B = np.asfortranarray(rng.standard_normal((n, 3)))
    got = po.csrmm("S", n, 3, 1.0, rp, ci, v, B, 0.0, np.zeros((n, 3)))
    want = po.ref_csrsymm1(n, 3, 1.0, rp + 1, ci + 1, v, B, np.zeros((n, 3), order="F"))
    assert np.array_equal(got, want)


@needs_ref
def test_sp_norm_f_vs_reference(bfwb398, pde2961):
    for (n, rp, ci, v, sym) in (bfwb398, pde2961):
        assert po.sp_norm_f(n, rp, ci, v, sym) == po.ref_sp_norm_f(n, rp, ci, v, sym)


def test_sp_norm_f_dense(bfwb398):
    n, rp, ci, v, _ = bfwb398
    A = expand_lower(n, rp, ci, v).toarray()
    assert abs(po.sp_norm_f(n, rp, ci, v, True) - np.linalg.norm(A)) <= 1e-13 * np.linalg.norm(A)


def test_gemv_vs_scipy_and_dense(pde2961):
    n, rp, ci, v, _ = pde2961
    A = to_scipy(n, rp, ci, v)
    x = np.random.default_rng(6).standard_normal(n)
    for t, M in (("n", A), ("t", A.T)):
        y = po.csrgemv(t, n, rp, ci, v, x)
        want = M @ x
        assert np.linalg.norm(y - want) <= 1e-13 * np.linalg.norm(want)
    n, rp, ci, v = rand_csr(40, 6, 7)
    D = to_scipy(n, rp, ci, v).toarray()
    x = np.random.default_rng(8).standard_normal(n)
    assert np.allclose(po.csrgemv("n", n, rp, ci, v, x), D @ x, rtol=1e-14, atol=1e-14)
    assert np.allclose(po.csrgemv("t", n, rp, ci, v, x), D.T @ x, rtol=1e-14, atol=1e-14)


def test_symv_expands_triangle(bfwb398):
    n, rp, ci, v, _ = bfwb398
    A = expand_lower(n, rp, ci, v)
    x = np.random.default_rng(9).standard_normal(n)
    y = po.csrsymv(n, rp, ci, v, x)
    assert np.linalg.norm(y - A @ x) <= 1e-13 * np.linalg.norm(A @ x)


def laplace_eigvec(m, k):
    g = np.arange(1, m + 1)
    s = [np.sin(kk * np.pi * g / (m + 1)) for kk in k]
    v = np.einsum("i,j,k->kji", s[0], s[1], s[2]).ravel()   # x fastest
    lam = sum(2 - 2 * np.cos(kk * np.pi / (m + 1)) for kk in k)
    return v, lam


def test_poisson_closed_form():
    from bksp_amd import gen
    m = 12
    n, rp, ci, v = gen.poisson3d(m, lower=True)
    assert rp[-1] == n + 3 * (m - 1) * m * m
    for k in ((1, 1, 1), (3, 5, 2), (12, 1, 7)):
        x, lam = laplace_eigvec(m, k)
        y = po.csrsymv(n, rp, ci, v, x)
        assert np.linalg.norm(y - lam * x) <= 1e-13 * lam * np.linalg.norm(x)


def test_convdiff2d_closed_form():
    """cfg1's operator: centred convection-diffusion is diagonally similar to a
    symmetric tridiagonal sum, so its eigenvalues are real and known."""
    from bksp_amd import gen
    m = 16
    n, rp, ci, v = gen.convdiff2d(m, 10.0, 20.0)
    assert rp[-1] == 5 * n - 4 * m
    A = to_scipy(n, rp, ci, v).toarray()
    h = 1.0 / (m + 1)
    ax = np.sqrt((1 / h**2) ** 2 - (10 / (2 * h)) ** 2)
    ay = np.sqrt((1 / h**2) ** 2 - (20 / (2 * h)) ** 2)
    k = np.arange(1, m + 1)
    lam = np.sort((2 / h**2 - 2 * ax * np.cos(k * np.pi / (m + 1)))[:, None]
                  + (2 / h**2 - 2 * ay * np.cos(k * np.pi / (m + 1)))[None, :], axis=None)
    ev = np.sort(np.linalg.eigvals(A).real)
    assert np.allclose(ev, lam, rtol=1e-9)


def test_dhorner_closed_form():
    rng = np.random.default_rng(10)
    for l in (1, 2, 4, 6):
        gam = rng.standard_normal(l)
        sig = float(rng.uniform(0.01, 1.0))
        out = po.dhorner(gam, l, sig)
        # q(lam) = 1 - sum_i gam[i-1] lam^i  ==  psi * (1 - sum_i out[i] (lam + sig)^i)
        for lam in rng.standard_normal(5):
            q = 1 - sum(gam[i - 1] * lam**i for i in range(1, l + 1))
            r = out[0] * (1 - sum(out[i] * (lam + sig) ** i for i in range(1, l + 1)))
            assert abs(q - r) <= 1e-12 * (1 + abs(q))


def test_comb_pascal():
    for n in range(0, 12):
        for r in range(-1, n + 2):
            assert po.comb(n, r) == (math.comb(n, r) if 0 <= r <= n else 0)


def test_myqr_invariants_and_lapack_signs():
    rng = np.random.default_rng(11)
    V = rng.standard_normal((300, 8))
    Q, R = po.myqr(V)
    assert np.allclose(Q.T @ Q, np.eye(8), atol=1e-14)
    assert np.allclose(Q @ R, V, atol=1e-13)
    assert np.allclose(np.tril(R, -1), 0)
    Qn, Rn = np.linalg.qr(V)          # LAPACK dgeqrf/dorgqr: same sign convention
    assert np.allclose(R, Rn, atol=1e-12) and np.allclose(Q, Qn, atol=1e-12)


def test_blas_semantics():
    rng = np.random.default_rng(12)
    x = rng.standard_normal(1000)
    assert abs(po.dnrm2(x) - np.linalg.norm(x)) <= 1e-15 * np.linalg.norm(x) * 4
    assert abs(po.dnrm2(x * 1e200) - np.linalg.norm(x) * 1e200) <= 1e-14 * np.linalg.norm(x) * 1e200
    y = rng.standard_normal(1000)
    assert po.ddot(x, y) == pytest.approx(float(x @ y), rel=1e-13)
