"""GPU parity of the CSR-stream SpMV (K1) against the oracle's restatement of
src/matvec.c. Stream-mode rows are summed in the reference's order with
separate multiply and add, so the expected result is BITWISE equal; rows longer
than the LDS capacity use a tree reduction and are held to 1e-13 relative."""
import numpy as np
import pytest

from bksp_amd import device as dv
from bksp_amd import gen
from conftest import expand_lower
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def gpu_spmv(n, rp, ci, v, x, symmetric=False, op="n", alpha=1.0, beta=0.0, y0=None):
    M = dv.Matrix(n, rp, ci, v, symmetric=symmetric, transpose=(op == "t"))
    dx = dv.DeviceArray.from_numpy(np.asarray(x, np.float64))
    dy = dv.DeviceArray.from_numpy(np.zeros(n) if y0 is None else np.asarray(y0, np.float64))
    M.spmv(dx, dy, op, alpha, beta)
    return dy.numpy()


def test_gemv_bitwise_fixture(pde2961):
    n, rp, ci, v, _ = pde2961
    x = np.random.default_rng(1).standard_normal(n)
    for op in "nt":
        assert np.array_equal(gpu_spmv(n, rp, ci, v, x, op=op), po.csrgemv(op, n, rp, ci, v, x))


def test_symv_bitwise_fixture(bfwb398):
    n, rp, ci, v, sym = bfwb398
    x = np.random.default_rng(2).standard_normal(n)
    assert np.array_equal(gpu_spmv(n, rp, ci, v, x, symmetric=True), po.csrsymv(n, rp, ci, v, x))


def test_alpha_beta_bitwise(pde2961):
    """y = beta*y + (alpha*a)*x in the order of src/matvec.c:140-143 (dcsrgemm1)."""
    n, rp, ci, v, _ = pde2961
    rng = np.random.default_rng(3)
    x, y0 = rng.standard_normal(n), rng.standard_normal(n)
    got = gpu_spmv(n, rp, ci, v, x, alpha=-0.75, beta=1.25, y0=y0)
    want = po.csrmm("G", n, 1, -0.75, rp, ci, v, x[:, None], 1.25, y0[:, None])[:, 0]
    assert np.array_equal(got, want)


def test_poisson_128_full_size_bitwise_and_closed_form():
    """cfg2's matrix at full size: 2,097,152 rows, 8,339,456 stored / 14,581,760 expanded."""
    m = 128
    n, rp, ci, v = gen.poisson3d(m, lower=True)
    assert rp[-1] == 8339456
    M = dv.Matrix(n, rp, ci, v, symmetric=True)
    assert M.nnz == 14581760
    x = np.random.default_rng(4).standard_normal(n)
    dx, dy = dv.DeviceArray.from_numpy(x), dv.DeviceArray((n,))
    M.spmv(dx, dy)
    assert np.array_equal(dy.numpy(), po.csrsymv(n, rp, ci, v, x))
    g = np.arange(1, m + 1)
    k = (3, 7, 11)
    s = [np.sin(kk * np.pi * g / (m + 1)) for kk in k]
    ev = np.einsum("i,j,k->kji", s[0], s[1], s[2]).ravel()
    lam = sum(2 - 2 * np.cos(kk * np.pi / (m + 1)) for kk in k)
    dx.upload(ev)
    M.spmv(dx, dy)
    assert np.linalg.norm(dy.numpy() - lam * ev) <= 1e-13 * lam * np.linalg.norm(ev)


def test_long_rows_and_empty_rows():
    rng = np.random.default_rng(5)
    n = 6000
    lens = rng.integers(0, 12, n)
    lens[[7, 3000, 5999]] = [2500, 5000, 4097]       # longer than the 2048-slot LDS stage
    lens[[10, 11, 12]] = 0                             # empty rows
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    ci = np.concatenate([rng.choice(n, size=l, replace=False) for l in lens]).astype(np.int32)
    v = rng.standard_normal(rp[-1])
    x = rng.standard_normal(n)
    got = gpu_spmv(n, rp, ci, v, x)
    want = po.csrgemv("n", n, rp, ci, v, x)
    absrow = np.add.reduceat(np.abs(v * x[ci]), rp[:-1]) * (lens > 0)
    assert np.all(np.abs(got - want) <= 1e-13 * absrow + 0.0)
    short = lens <= 2048
    assert np.array_equal(got[short], want[short])
    assert np.all(got[[10, 11, 12]] == 0)


def test_fp32_spmv_close():
    n, rp, ci, v = gen.convdiff3d(24, *gen.cfg3_peclet())
    M = dv.Matrix(n, rp, ci, v, fp32=True)
    x = np.random.default_rng(6).standard_normal(n)
    dx, dy = dv.DeviceArray.from_numpy(x.astype(np.float32)), dv.DeviceArray((n,), np.float32)
    M.spmv32(dx, dy)
    want = po.csrgemv("n", n, rp, ci, v, x.astype(np.float32).astype(np.float64))
    absrow = np.add.reduceat(np.abs(v * x[ci]), rp[:-1])
    assert np.all(np.abs(dy.numpy() - want) <= 2e-6 * absrow)


def test_symmetric_expansion_matches_scipy(bfwb398):
    n, rp, ci, v, _ = bfwb398
    A = expand_lower(n, rp, ci, v)
    x = np.random.default_rng(7).standard_normal(n)
    got = gpu_spmv(n, rp, ci, v, x, symmetric=True)
    assert np.linalg.norm(got - A @ x) <= 1e-13 * np.linalg.norm(A @ x)


def test_dot_deterministic():
    rng = np.random.default_rng(8)
    x, y = rng.standard_normal(3_000_001), rng.standard_normal(3_000_001)
    dx, dy = dv.DeviceArray.from_numpy(x), dv.DeviceArray.from_numpy(y)
    a, b = dv.dot(dx, dy), dv.dot(dx, dy)
    assert a == b
    assert abs(a - float(x @ y)) <= 1e-12 * np.abs(x * y).sum()


@pytest.mark.parametrize("variant", [0, 1, 2, 3])
def test_every_kernel_variant_bitwise(pde2961, bfwb398, variant):
    """Workgroup CSR-stream (0), wave CSR-stream (1), SELL-64 with 16-bit column deltas
    (2) and with 32-bit columns (3): every variant sums rows in stored order."""
    cases = [(pde2961[:4], False), (bfwb398[:4], True), (gen.poisson3d(24, lower=True), True)]
    for (n, rp, ci, v), sym in cases:
        M = dv.Matrix(n, rp, ci, v, symmetric=sym, transpose=not sym)
        try:
            M.set_spmv(variant)
        except dv.BkspDeviceError:
            assert variant >= 2            # SELL needs small padding; CSR-stream always exists
            continue
        x = np.random.default_rng(5).standard_normal(n)
        dx, dy = dv.DeviceArray.from_numpy(x), dv.DeviceArray.from_numpy(np.zeros(n))
        for op in ("n",) if sym else ("n", "t"):
            M.spmv(dx, dy, op)
            ref = po.csrsymv(n, rp, ci, v, x) if sym else po.csrgemv(op, n, rp, ci, v, x)
            assert np.array_equal(dy.numpy(), ref), (n, op, variant)
            # alpha/beta form
            y0 = np.random.default_rng(6).standard_normal(n)
            dy.upload(y0)
            M.spmv(dx, dy, op, 0.5, 2.0)
            assert np.allclose(dy.numpy(), 2.0 * y0 + 0.5 * ref, rtol=1e-13, atol=1e-13)
