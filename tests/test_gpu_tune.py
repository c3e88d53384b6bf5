"""The matrix autotuner (bkh_matrix_tune): whatever SpMV variant, wave blocking and
SpMM columns per pass it keeps, every product stays bitwise the oracle's restatement of
src/matvec.c (each row sums its entries in stored order), and the solvers run on the
tuned matrix with unchanged results."""
import numpy as np
import pytest

from bksp_amd import device as dv
from bksp_amd import gen
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def gpu_spmm(M, n, s, X, alpha=1.0, beta=0.0, Y0=None):
    dX = dv.DeviceArray.from_numpy(np.asfortranarray(X).ravel(order="F"))
    dY = dv.DeviceArray.from_numpy(np.zeros(n * s) if Y0 is None else np.asfortranarray(Y0).ravel(order="F"))
    M.spmm(s, dX, dY, alpha, beta)
    return dY.numpy().reshape(s, n).T


def check_products(M, n, rp, ci, v, sym, seed):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(n)
    dx, dy = dv.DeviceArray.from_numpy(x), dv.DeviceArray.from_numpy(np.zeros(n))
    M.spmv(dx, dy)
    ref = po.csrsymv(n, rp, ci, v, x) if sym else po.csrgemv("n", n, rp, ci, v, x)
    assert np.array_equal(dy.numpy(), ref)
    kind = "S" if sym else "G"
    for s in (1, 3, 8, 16, 20, 37):
        X, Y0 = rng.standard_normal((n, s)), rng.standard_normal((n, s))
        assert np.array_equal(gpu_spmm(M, n, s, X), po.csrmm(kind, n, s, 1.0, rp, ci, v, X, 0.0, np.zeros((n, s)))), s
        got = gpu_spmm(M, n, s, X, -0.75, 1.25, Y0)
        assert np.array_equal(got, po.csrmm(kind, n, s, -0.75, rp, ci, v, X, 1.25, Y0)), s


@pytest.mark.parametrize("case", ["poisson", "pde2961", "random_rows", "bfwb398"])
def test_tuned_matrix_products_bitwise(case, pde2961, bfwb398):
    if case == "poisson":
        (n, rp, ci, v), sym = gen.poisson3d(24, lower=True), True
    elif case == "pde2961":
        (n, rp, ci, v), sym = pde2961[:4], False
    elif case == "bfwb398":
        (n, rp, ci, v), sym = bfwb398[:4], True
    else:
        (n, rp, ci, v), sym = gen.random_rows(3131, 5000), False
    M = dv.Matrix(n, rp, ci, v, symmetric=sym, fp32=True)
    check_products(M, n, rp, ci, v, sym, 11)           # before tuning (s > 16 in chunks of 16)
    r = M.tune(16)
    print(case, {k: r[k] for k in ("spmv_variant", "rows_per_wave", "nnz_per_wave", "candidates",
                                   "spmv_ms", "spmv_ms_default")}, "spmm widths", r["spmm_width"])
    assert r["spmv_variant"] in (0, 1, 2, 3) and r["candidates"] >= 1
    # separate timings of microsecond kernels: only a gross regression fails
    assert 0 < r["spmv_ms"] <= r["spmv_ms_default"] * 3 + 0.01
    for s in range(1, 17):
        assert 1 <= r["spmm_width"][s] <= s
        assert r["spmm_ms"][s] <= r["spmm_ms_full"][s]
    check_products(M, n, rp, ci, v, sym, 12)
    # each forced wave blocking of the candidates is bitwise too: force the wave
    # variant after tuning when the matrix allows it
    if max(np.diff(rp)) <= 512:
        M.set_spmv(1)
        check_products(M, n, rp, ci, v, sym, 13)


def test_solvers_on_tuned_matrix():
    """CG (FP64 and FP32-IR) on a tuned 32^3 Poisson matrix: the same iterate and
    solution as on the untuned matrix."""
    n, rp, ci, v = gen.poisson3d(32, lower=True)
    b = gen.normal(77, n)
    out = []
    for tune in (False, True):
        M = dv.Matrix(n, rp, ci, v, symmetric=True, fp32=True)
        if tune:
            M.tune(8)
        res = []
        for prec in (dv.BKH_FP64, dv.BKH_FP32_IR):
            dx = dv.DeviceArray.from_numpy(np.zeros(n))
            info = dv.cg(M, dv.DeviceArray.from_numpy(b), dx, 1e-8, 2000, precision=prec)
            res.append((info.ret, dx.numpy()))
        out.append(res)
    for (ra, xa), (rb, xb) in zip(out[0], out[1]):
        assert ra == rb
        assert np.array_equal(xa, xb)
