"""Pins the oracle's complex FP64 family (oracle/orc_complex.c) before it judges
the GPU: the CSR kernels bitwise against the reference's own src/matvec.c
(compiled unmodified under oracle/_ref) and against scipy; the solvers against
the real-valued oracle on real data (the complex recurrences reduce to the real
ones when every imaginary part is zero), against the true solution of seeded
complex systems, and every shifted solution against its own true residual."""
import numpy as np
import pytest

from conftest import expand_lower, zcsr_dense, zsystem
from oracle import pyoracle as po


def zrand(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


@pytest.mark.parametrize("trans", ["n", "t"])
def test_zcsrgemm0_against_reference_and_scipy(trans):
    n, rp, ci, v = zsystem(24, seed=1)
    rng = np.random.default_rng(2)
    X = zrand(rng, n, 3)
    got = po.zcsrgemm0(trans, n, rp, ci, v, X)
    A = zcsr_dense(n, rp, ci, v)
    want = (A if trans == "n" else A.conj().T) @ X
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()
    if po.ref() is not None:
        assert np.array_equal(got, po.ref_zcsrgemm0(trans, n, rp, ci, v, X))


@pytest.mark.parametrize("trans", ["n", "t"])
def test_zcsrgemm1_against_reference(trans):
    n, rp, ci, v = zsystem(16, seed=3)
    rng = np.random.default_rng(4)
    X, C = zrand(rng, n, 2), zrand(rng, n, 2)
    al, be = 0.7 - 0.4j, -1.25 + 0.5j
    got = po.zcsrgemm1(trans, al, n, rp, ci, v, X, be, C)
    A = zcsr_dense(n, rp, ci, v)
    want = be * C + al * ((A if trans == "n" else A.conj().T) @ X)
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()
    if po.ref() is not None:
        assert np.array_equal(got, po.ref_zcsrgemm1(trans, al, n, rp, ci, v, X, be, C))


def test_level1_semantics():
    rng = np.random.default_rng(5)
    x, y = zrand(rng, 1000), zrand(rng, 1000)
    assert abs(po.zdotc(x, y) - np.vdot(x, y)) <= 1e-12 * abs(np.vdot(x, y))
    assert abs(po.dznrm2(x) - np.linalg.norm(x)) <= 1e-14 * np.linalg.norm(x)
    big = x * 1e300                                   # the scaled sum does not overflow
    assert np.isfinite(po.dznrm2(big)) and abs(po.dznrm2(big) / 1e300 - np.linalg.norm(x)) <= 1e-13 * np.linalg.norm(x)


def test_zhorner_closed_form_and_real_case():
    rng = np.random.default_rng(6)
    for l in (1, 2, 4, 6):
        gam = zrand(rng, l)
        sig = complex(rng.uniform(0.01, 1.0), rng.uniform(-1.0, 1.0))
        out = po.zhorner(gam, l, sig)
        # q(t) = 1 - sum_i gam[i-1] t^i  ==  out0 * (1 - sum_i out[i] (t + sig)^i)
        for t in zrand(rng, 5):
            q = 1 - sum(gam[i - 1] * t**i for i in range(1, l + 1))
            r = out[0] * (1 - sum(out[i] * (t + sig) ** i for i in range(1, l + 1)))
            assert abs(q - r) <= 1e-11 * (1 + abs(q))
        gr, sr = rng.standard_normal(l), float(rng.uniform(0.01, 1.0))
        assert np.array_equal(po.zhorner(gr.astype(complex), l, sr).real, po.dhorner(gr, l, sr))


def real_system():
    from bksp_amd import gen
    n, rp, ci, v = gen.convdiff2d(20, 10.0, 20.0)
    return n, rp.astype(np.int32), ci.astype(np.int32), v


def test_real_data_reduces_to_real_solvers():
    """With every imaginary part zero, each complex recurrence performs the real one's
    arithmetic (zero products and sums are exact), so iterates agree with the
    real oracle: bitwise for BiCG and BiCGSTAB(l); the shifted solver differs only
    by cpow(sigma, k) against pow(sigma, k) in its binomial matrix."""
    n, rp, ci, v = real_system()
    b = np.random.default_rng(7).standard_normal(n)
    z = v.astype(complex)
    ret, x = po.bicg(n, rp, ci, v, b, 1e-10, 500)
    zret, zx = po.zbicg(n, rp, ci, z, b.astype(complex), 1e-10, 500)
    assert ret == zret and ret > 0 and np.array_equal(zx.real, x) and not zx.imag.any()
    for l in (1, 2, 4):
        ret, x = po.bicgstabl(n, rp, ci, v, b, l, 1e-10, 500)
        zret, zx = po.zbicgstabl(n, rp, ci, z, b.astype(complex), l, 1e-10, 500)
        assert ret == zret and ret > 0 and np.array_equal(zx.real, x), l
    sig = np.array([0.05, 0.3, 1.0])
    ret, X = po.shbicgstabl(n, rp, ci, v, b, 3, sig, 1e-10, 500)
    zret, ZX = po.zshbicgstabl(n, rp, ci, z, b.astype(complex), 3, sig.astype(complex), 1e-10, 500)
    assert ret == zret and ret > 0
    assert np.array_equal(ZX[0].real, X[0])
    assert np.abs(ZX.real - X).max() <= 1e-9 * np.abs(X).max() and np.abs(ZX.imag).max() <= 1e-12 * np.abs(X).max()


def true_res(A, x, b, shift=0.0):
    return np.linalg.norm(b - A @ x - shift * x) / np.linalg.norm(b)


def test_complex_solvers_reach_the_true_solution():
    n, rp, ci, v = zsystem(24, seed=8)
    A = zcsr_dense(n, rp, ci, v)
    rng = np.random.default_rng(9)
    xt = zrand(rng, n)
    b = A @ xt
    ret, x = po.zbicg(n, rp, ci, v, b, 1e-10, 2000)
    assert ret >= 0 and true_res(A, x, b) < 1e-9 and np.linalg.norm(x - xt) < 1e-6 * np.linalg.norm(xt)
    for l in (1, 2, 4):
        ret, x = po.zbicgstabl(n, rp, ci, v, b, l, 1e-10, 2000)
        assert ret >= 0 and true_res(A, x, b) < 1e-9, (l, ret)


def test_zbicg_diagonal_closed_form():
    """On a diagonal matrix with k distinct eigenvalues BiCG terminates (to rounding)
    after k steps: the breaking index is k - 1."""
    d = np.repeat(np.array([1.0 + 1j, 2.0 - 0.5j, 3.0 + 0.25j, 5.0]), 10)
    n = len(d)
    rp, ci = np.arange(n + 1, dtype=np.int32), np.arange(n, dtype=np.int32)
    b = zrand(np.random.default_rng(10), n)
    ret, x = po.zbicg(n, rp, ci, d, b, 1e-12, 50)
    assert ret == 3 and np.abs(x - b / d).max() <= 1e-12 * np.abs(b / d).max()


def test_zshbicgstabl_every_shift_solves_its_system():
    n, rp, ci, v = zsystem(20, seed=11)
    A = zcsr_dense(n, rp, ci, v)
    b = zrand(np.random.default_rng(12), n)
    sig = np.array([0.02 + 0.1j, 0.5 - 0.2j, 1.0, 2.0 + 1.0j])
    for l in (1, 2, 4):
        ret, X = po.zshbicgstabl(n, rp, ci, v, b, l, sig, 1e-10, 1000)
        assert ret >= 0
        assert true_res(A, X[0], b) < 1e-10
        for k, s in enumerate(sig):
            assert true_res(A, X[k + 1], b, s) < 1e-8, (l, k, true_res(A, X[k + 1], b, s))


def test_zqreco_invariants():
    rng = np.random.default_rng(13)
    V = zrand(rng, 300, 6)
    Q, R = po.zqreco(V)
    assert np.allclose(Q.conj().T @ Q, np.eye(6), atol=1e-14)
    assert np.allclose(Q @ R, V, atol=1e-13)
    assert np.allclose(np.tril(R, -1), 0) and np.abs(np.diag(R).imag).max() == 0.0   # zgeqrf: real diagonal
    Qn, Rn = np.linalg.qr(V)          # LAPACK zgeqrf/zungqr: the same reflector convention
    assert np.allclose(R, Rn, atol=1e-12) and np.allclose(Q, Qn, atol=1e-12)


def test_zblbicgstabl_real_data_and_true_residual():
    """On real data the complex block solver performs the real block solver's
    operations (zgemm 'T' is dgemm 'T', zqreco is myqr); on a complex system every
    column reaches the true residual."""
    n, rp, ci, v = real_system()
    B = np.random.default_rng(14).standard_normal((n, 3))
    ret, X, tr = po.blbicgstabl(n, rp, ci, v, B, 2, 1e-10, 200)
    zret, ZX, ztr = po.zblbicgstabl(n, rp, ci, v.astype(complex), B.astype(complex), 2, 1e-10, 200)
    assert ret == zret and ret > 0 and np.abs(ZX.real - X).max() <= 1e-10 * np.abs(X).max()
    n, rp, ci, v = zsystem(20, seed=15)
    A = zcsr_dense(n, rp, ci, v)
    B = zrand(np.random.default_rng(16), n, 4)
    for l in (1, 2, 4):
        ret, X, tr = po.zblbicgstabl(n, rp, ci, v, B, l, 1e-10, 500)
        res = np.linalg.norm(B - A @ X) / np.linalg.norm(B)
        assert ret >= 0 and res < 1e-8 and abs(tr - res) <= 1e-6 * res, (l, ret, res, tr)
