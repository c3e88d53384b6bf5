"""Pins the oracle's solvers (CPU only) with what the mathematics fixes.

* CG after k steps is the Galerkin solution on K_k(A, r0); CR the minimum-residual
  one; BiCG the Petrov-Galerkin one with test space K_k(A^T, r0) (dense closed forms).
* CG on a matrix with k distinct eigenvalues stops at loop index k - 1.
* One outer BiCGSTAB(l) iteration = l BiCG steps + the degree-l minimum-residual
  polynomial (closed form).
* Shifted BiCGSTAB(l): the seed iterate equals BiCGSTAB(l)'s, each shifted x solves
  (A + sigma I) x = b.
* Block CG's first step with the reference's P0 = 2V is half the block Galerkin
  step (P0 = V gives the full step).
* Block BiCGSTAB(l) with one column reproduces BiCGSTAB(l) iterate by iterate.
* Every solve returns x with a small true residual on the reference's fixtures.
A dropped term, a wrong sign or index, or a transposed operand breaks one of these.
"""
import numpy as np
import pytest

from conftest import expand_lower, to_scipy
from oracle import pyoracle as po

TINY = 1e-300


def spd_lower(n, seed, dens=0.2):
    rng = np.random.default_rng(seed)
    M = rng.standard_normal((n, n)) * (rng.random((n, n)) < dens)
    A = M @ M.T + n * np.eye(n)
    L = np.tril(A)
    rp = np.concatenate([[0], np.cumsum((L != 0).sum(1))]).astype(np.int32)
    ci = np.concatenate([np.nonzero(L[i])[0] for i in range(n)]).astype(np.int32)
    v = np.concatenate([L[i][L[i] != 0] for i in range(n)])
    return A, rp, ci, v


def gen_dense(n, seed, shift=3.0):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, n)) * (rng.random((n, n)) < 0.3) + shift * np.sqrt(n) * np.eye(n)
    rp = np.concatenate([[0], np.cumsum((A != 0).sum(1))]).astype(np.int32)
    ci = np.concatenate([np.nonzero(A[i])[0] for i in range(n)]).astype(np.int32)
    v = np.concatenate([A[i][A[i] != 0] for i in range(n)])
    return A, rp, ci, v


def krylov(A, r, k):
    K = np.empty((len(r), k))
    K[:, 0] = r / np.linalg.norm(r)
    for j in range(1, k):
        w = A @ K[:, j - 1]
        K[:, j] = w / np.linalg.norm(w)
    return np.linalg.qr(K)[0]


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("k", [1, 2, 4, 6])
def test_cg_galerkin_closed_form(k):
    n = 30
    A, rp, ci, v = spd_lower(n, 20 + k)
    b = np.random.default_rng(k).standard_normal(n)
    ret, x = po.cg(n, rp, ci, v, b, TINY, k)
    assert ret == -1
    V = krylov(A, b, k)
    want = V @ np.linalg.solve(V.T @ A @ V, V.T @ b)
    assert rel(x, want) < 1e-9


def test_cg_distinct_eigenvalues_index():
    lam = np.array([1.0, 2.0, 3.0, 5.0])
    d = np.tile(lam, 10)
    n = len(d)
    rp = np.arange(n + 1, dtype=np.int32)
    ci = np.arange(n, dtype=np.int32)
    ret, x = po.cg(n, rp, ci, d, np.ones(n), 1e-10, 100)
    assert ret == 3
    assert rel(x, 1.0 / d) < 1e-12


@pytest.mark.parametrize("k", [1, 3, 5])
def test_cr_minres_closed_form(k):
    n = 30
    A, rp, ci, v = spd_lower(n, 40 + k)
    b = np.random.default_rng(k + 7).standard_normal(n)
    ret, x = po.cr(n, rp, ci, v, b, TINY, k + 1)    # CR's loop starts at i = 1
    assert ret == -1
    V = krylov(A, b, k)
    y = np.linalg.lstsq(A @ V, b, rcond=None)[0]
    assert rel(x, V @ y) < 1e-9


@pytest.mark.parametrize("k", [1, 2, 4])
def test_bicg_petrov_galerkin_closed_form(k):
    n = 30
    A, rp, ci, v = gen_dense(n, 60 + k)
    b = np.random.default_rng(k + 9).standard_normal(n)
    ret, x = po.bicg(n, rp, ci, v, b, TINY, k)
    assert ret == -1
    V, W = krylov(A, b, k), krylov(A.T, b, k)
    want = V @ np.linalg.solve(W.T @ A @ V, W.T @ b)
    assert rel(x, want) < 1e-8


@pytest.mark.parametrize("l", [1, 2, 4])
def test_bicgstabl_one_outer_closed_form(l):
    n = 40
    A, rp, ci, v = gen_dense(n, 80 + l)
    b = np.random.default_rng(l + 11).standard_normal(n)
    ret, x = po.bicgstabl(n, rp, ci, v, b, l, TINY, 1)
    V, W = krylov(A, b, l), krylov(A.T, b, l)
    xb = V @ np.linalg.solve(W.T @ A @ V, W.T @ b)
    rb = b - A @ xb
    R = [rb]
    for j in range(l):
        R.append(A @ R[-1])
    g = np.linalg.lstsq(np.stack(R[1:], 1), rb, rcond=None)[0]
    want = xb + np.stack(R[:-1], 1) @ g
    assert rel(x, want) < 1e-8


def test_shbicgstabl_seed_equals_bicgstabl_and_shifts_solve(pde2961):
    n, rp, ci, v, _ = pde2961
    A = to_scipy(n, rp, ci, v)
    b = np.zeros(n); b[0] = 1.0
    sig = [0.0, 0.2, 0.4, 0.8]
    for l in (1, 2, 4):
        # same iterates (the two files order the final r0 update differently, so
        # they agree to rounding, which the erratic BiCGSTAB history amplifies)
        _, x1 = po.bicgstabl(n, rp, ci, v, b, l, TINY, 10)
        _, X = po.shbicgstabl(n, rp, ci, v, b, l, sig, TINY, 10)
        assert rel(X[0], x1) < 1e-12
        r1, x1 = po.bicgstabl(n, rp, ci, v, b, l, 1e-10, 300)
        r2, X = po.shbicgstabl(n, rp, ci, v, b, l, sig, 1e-10, 300)
        assert r1 >= 0 and abs(r1 - r2) <= max(2, 0.1 * r1)
        for k, s in enumerate(sig):
            res = np.linalg.norm(b - A @ X[k + 1] - s * X[k + 1]) / np.linalg.norm(b)
            assert res < 1e-8, (l, s, res)
        assert rel(X[1], X[0]) < 1e-8          # the sigma = 0 shift is the seed system


def test_blcg_first_step_closed_form(bfwb398):
    n, rp, ci, v, _ = bfwb398
    A = expand_lower(n, rp, ci, v).toarray()
    rng = np.random.default_rng(13)
    B = np.asfortranarray(rng.standard_normal((n, 4)))
    V = np.linalg.qr(B)[0]
    full = V @ np.linalg.solve(V.T @ A @ V, V.T @ B)
    _, X = po.blcg(n, rp, ci, v, B, TINY, 1, p0_mode=0)
    assert rel(X, 0.5 * full) < 1e-10               # reference: P0 = 2V -> half step
    _, X = po.blcg(n, rp, ci, v, B, TINY, 1, p0_mode=1)
    assert rel(X, full) < 1e-10


def test_blcg_converges_both_readings():
    from bksp_amd import gen
    n, rp, ci, v = gen.poisson3d(10, lower=True)
    A = expand_lower(n, rp, ci, v)
    B = np.asfortranarray(np.random.default_rng(14).standard_normal((n, 4)))
    its = []
    for mode in (0, 1):
        ret, X = po.blcg(n, rp, ci, v, B, 1e-10, 500, p0_mode=mode)
        assert ret >= 0
        assert np.linalg.norm(B - A @ X) / np.linalg.norm(B) < 1e-8
        its.append(ret)
    assert its[0] > its[1]                           # the 2V start costs iterations


@pytest.mark.parametrize("l", [1, 2, 3])
def test_blbicgstabl_one_column_is_bicgstabl(pde2961, l):
    n, rp, ci, v, _ = pde2961
    b = np.random.default_rng(15).standard_normal(n)
    for k in (1, 3):
        _, x = po.bicgstabl(n, rp, ci, v, b, l, TINY, k)
        _, X, _ = po.blbicgstabl(n, rp, ci, v, b[:, None], l, TINY, k)
        assert rel(X[:, 0], x) < 1e-9


def test_reference_scenarios_true_residual(bfwb398, pde2961):
    """The scenarios of the reference's tests/test.c, with assertions."""
    n, rp, ci, v, _ = bfwb398
    A = expand_lower(n, rp, ci, v)
    b = np.zeros(n); b[0] = 1
    for fn in (po.cg, po.cr):
        ret, x = fn(n, rp, ci, v, b, 1e-14, 500)
        res = np.linalg.norm(b - A @ x)
        assert (ret >= 0 and res < 1e-10) or (ret == -1 and res < 1e-8), (fn.__name__, ret, res)
    B = np.zeros((n, 16), order="F"); B[np.arange(16), np.arange(16)] = 1
    ret, X = po.blcg(n, rp, ci, v, B, 1e-14, 500)
    assert np.linalg.norm(B - A @ X) / 4 < 1e-8
    n, rp, ci, v, _ = pde2961
    A = to_scipy(n, rp, ci, v)
    b = np.zeros(n); b[0] = 1
    for ret, x in (po.bicg(n, rp, ci, v, b, 1e-14, 500), po.bicgstabl(n, rp, ci, v, b, 4, 1e-14, 500)):
        assert np.linalg.norm(b - A @ x) < 1e-9


# ---- link-parity solvers: src/bicgstab.c, src/shbicg.c, src/blbicgstab.c


@pytest.mark.parametrize("k", [1, 2, 3, 6])
def test_bicgstab_equals_bicgstabl_one(k):
    """BiCGSTAB is BiCGSTAB(1) (van der Vorst 1992; Sleijpen-Fokkema 1993): the same
    iterates in exact arithmetic, through two different recurrences (bicgstab.c's
    rho/omega form, bicgstabl.c's BiCG step + degree-1 MR polynomial, itself pinned by
    its closed form above)."""
    n = 40
    A, rp, ci, v = gen_dense(n, 90 + k)
    b = np.random.default_rng(k + 13).standard_normal(n)
    ret, x = po.bicgstab(n, rp, ci, v, b, TINY, k)
    assert ret == -1
    _, xl = po.bicgstabl(n, rp, ci, v, b, 1, TINY, k)
    assert rel(x, xl) < 1e-9


def shifted_csr(n, rp, ci, v, sigma):
    """A + sigma I as CSR (scipy; the diagonal entry added or inserted)."""
    import scipy.sparse as sp
    S = (to_scipy(n, rp, ci, v) + sigma * sp.identity(n, format="csr")).tocsr()
    S.sort_indices()
    return S.indptr.astype(np.int32), S.indices.astype(np.int32), S.data.copy()


@pytest.mark.parametrize("k", [1, 3, 8])
def test_shbicg_seed_is_bicg_and_shifts_are_bicg_on_shifted_systems(pde2961, k):
    """Shifted BiCG (Frommer 2003): with x0 = 0 the shifted residuals stay collinear
    with the seed's, so the seed iterate is BiCG's on A and shift k's iterate is BiCG's
    on A + sigma_k I, iteration by iteration (orc_bicg pinned by its Petrov-Galerkin
    closed form above). A wrong pi recurrence or shift scaling breaks the second."""
    n, rp, ci, v, _ = pde2961
    b = np.random.default_rng(5).standard_normal(n)
    sig = [0.0, 0.2, 0.4, 0.8]
    ret, X = po.shbicg(n, rp, ci, v, b, sig, TINY, k)
    assert ret == -1
    _, x0 = po.bicg(n, rp, ci, v, b, TINY, k)
    assert rel(X[0], x0) < 1e-10
    for q, s in enumerate(sig):
        rps, cis, vs = shifted_csr(n, rp, ci, v, s)
        _, xs = po.bicg(n, rps, cis, vs, b, TINY, k)
        assert rel(X[q + 1], xs) < 1e-8, (q, s, rel(X[q + 1], xs))


def test_shbicg_reference_scenario_solves_every_system(pde2961):
    """tests/test.c:243-262 scenario: each (A + sigma_k I) x_k = b to the tolerance's scale."""
    n, rp, ci, v, _ = pde2961
    A = to_scipy(n, rp, ci, v)
    b = np.zeros(n); b[0] = 1.0
    sig = [0.0, 0.2, 0.4, 0.8]
    ret, X = po.shbicg(n, rp, ci, v, b, sig, 1e-14, 500)
    assert ret > 0
    for q, s in enumerate([0.0] + sig):
        assert np.linalg.norm(b - A @ X[q] - s * X[q]) < 1e-12


@pytest.mark.parametrize("k", [1, 2, 5])
def test_blbicgstab_one_column_is_bicgstab(k):
    """With one right-hand side the block method is BiCGSTAB: alpha, beta are 1 x 1
    and the orthonormalisation of P only rescales it, which changes neither P alpha
    nor the next direction (both invariant under P -> P d)."""
    n = 40
    A, rp, ci, v = gen_dense(n, 110 + k)
    b = np.random.default_rng(k + 17).standard_normal(n)
    ret, X = po.blbicgstab(n, rp, ci, v, b[:, None], TINY, k)
    assert ret == -1
    _, x = po.bicgstab(n, rp, ci, v, b, TINY, k)
    assert rel(X[:, 0], x) < 1e-9


def test_blbicgstab_first_step_and_convergence():
    """First iterate: the block Petrov-Galerkin step X = P0 alpha with
    Rs^T (R0 - A P0 alpha) = 0 (P0 = Rs = R0 = B), then the Frobenius minimal-residual
    step omega = argmin ||S - omega A S||_F on S = R0 - A P0 alpha; and a converged
    block solve has a small true residual for every column."""
    n, s = 50, 3
    A, rp, ci, v = gen_dense(n, 130)
    B = np.random.default_rng(131).standard_normal((n, s))
    ret, X = po.blbicgstab(n, rp, ci, v, B, TINY, 1)
    alpha = np.linalg.solve(B.T @ (A @ B), B.T @ B)
    S = B - A @ B @ alpha
    AS = A @ S
    omega = np.sum(AS * S) / np.sum(AS * AS)
    assert rel(X, B @ alpha + omega * S) < 1e-10
    ret, X = po.blbicgstab(n, rp, ci, v, B, 1e-12, 200)
    assert ret >= 0
    assert np.linalg.norm(B - A @ X) / np.linalg.norm(B) < 1e-10
