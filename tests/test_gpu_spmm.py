"""GPU parity of the block SpMM (bkh_spmm) against the oracle's restatement of
dcsrgemm1 / dcsrsymm1 (src/matvec.c:130-172, :219-246; the mkl_dcsrmm dispatch of
include/bksp_internal.h:63-79). On the SELL-64 copy, s * 8 in {16, 32, 64, 128}
bytes runs with s * 8 / 16 lanes per row (each lane a 16-byte segment of the row),
other s with a lane per row. Per column every row sums its entries in the stored
order with separate multiply and add, so results are BITWISE equal for every s."""
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


@pytest.mark.parametrize("s", list(range(1, 17)))
def test_general_bitwise_every_width(pde2961, s):
    n, rp, ci, v, _ = pde2961
    M = dv.Matrix(n, rp, ci, v)
    rng = np.random.default_rng(100 + s)
    X, Y0 = rng.standard_normal((n, s)), rng.standard_normal((n, s))
    assert np.array_equal(gpu_spmm(M, n, s, X), po.csrmm("G", n, s, 1.0, rp, ci, v, X, 0.0, np.zeros((n, s))))
    got = gpu_spmm(M, n, s, X, -0.75, 1.25, Y0)
    assert np.array_equal(got, po.csrmm("G", n, s, -0.75, rp, ci, v, X, 1.25, Y0))


@pytest.mark.parametrize("s", [1, 2, 4, 8, 16])
def test_symmetric_bitwise(bfwb398, s):
    n, rp, ci, v, _ = bfwb398
    M = dv.Matrix(n, rp, ci, v, symmetric=True)
    X = np.random.default_rng(200 + s).standard_normal((n, s))
    assert np.array_equal(gpu_spmm(M, n, s, X), po.csrmm("S", n, s, 1.0, rp, ci, v, X, 0.0, np.zeros((n, s))))


@pytest.mark.parametrize("s", [4, 8, 16])
def test_poisson_sell_bitwise_and_columns(s):
    """40^3 Poisson stored as its lower triangle (SELL-64, 16-bit column deltas): the
    oracle's symmetric product, and every column equal to the single-vector SpMV."""
    n, rp, ci, v = gen.poisson3d(40, lower=True)
    M = dv.Matrix(n, rp, ci, v, symmetric=True)
    X = np.random.default_rng(300 + s).standard_normal((n, s))
    got = gpu_spmm(M, n, s, X)
    assert np.array_equal(got, po.csrmm("S", n, s, 1.0, rp, ci, v, X, 0.0, np.zeros((n, s))))
    dx, dy = dv.DeviceArray((n,)), dv.DeviceArray((n,))
    for j in (0, s - 1):
        dx.upload(np.ascontiguousarray(X[:, j]))
        M.spmv(dx, dy)
        assert np.array_equal(got[:, j], dy.numpy())
