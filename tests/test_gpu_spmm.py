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


def gpu_spmm_gram(M, n, s, P):
    dP = dv.DeviceArray.from_numpy(np.asfortranarray(P).ravel(order="F"))
    dAP = dv.DeviceArray((n * s,))
    G, path = M.spmm_gram(s, dP, dAP)
    return dAP.numpy().reshape(s, n).T, G, path


@pytest.mark.parametrize("s", [1, 2, 4, 8, 16])
@pytest.mark.parametrize("m", [24, 33])
def test_spmm_gram_pipelined_dia(m, s):
    """Block CG's AP = A P with P^T AP (src/blcg.c:74-79) on Poisson grids whose slices
    carry per-slice records (m = 33: n = 35,937 ends in a partial slice and the first
    and last planes need clamped gathers): AP is bitwise the oracle's symmetric product;
    the Gram agrees with P^T (AP) formed on the host in FP64 to 1e-13 relative."""
    n, rp, ci, v = gen.poisson3d(m, lower=True)
    M = dv.Matrix(n, rp, ci, v, symmetric=True)
    P = np.random.default_rng(400 + s + m).standard_normal((n, s))
    AP, G, path = gpu_spmm_gram(M, n, s, P)
    want = po.csrmm("S", n, s, 1.0, rp, ci, v, P, 0.0, np.zeros((n, s)))
    assert np.array_equal(AP, want)
    if s * 8 in (16, 32, 64, 128):
        assert path == 2
    ref = P.T @ want
    assert np.max(np.abs(G - ref)) <= 1e-13 * np.max(np.abs(ref))


def lower_of(n, rp, ci, v):
    """The lower triangle (diagonal included) of a CSR matrix, as the stored triangle of
    a symmetric one."""
    rows = np.repeat(np.arange(n), np.diff(rp))
    keep = ci <= rows
    lr = np.concatenate([[0], np.cumsum(np.bincount(rows[keep], minlength=n))]).astype(np.int32)
    return lr, ci[keep].astype(np.int32), v[keep]


@pytest.mark.parametrize("s", [2, 8])
@pytest.mark.parametrize("kind", ["bfwb398", "random"])
def test_spmm_gram_other_matrices(bfwb398, kind, s):
    """Symmetric matrices (block CG's input, src/blcg.c:18-20) without per-slice
    records: the reference's bfwb398 fixture and the lower triangle of a random-column
    matrix (no slice-diagonal copy: the SELL fused kernel or the separate SpMM and Gram
    pass, which mirrors the upper triangle of the symmetric P^T A P). AP bitwise, the
    Gram to 1e-13 relative of P^T (AP) on the host."""
    if kind == "bfwb398":
        n, rp, ci, v, _ = bfwb398
    else:
        n, rp0, ci0, v0 = gen.random_rows(606, 3000)
        rp, ci, v = lower_of(n, rp0, ci0, v0)
    M = dv.Matrix(n, rp, ci, v, symmetric=True)
    P = np.random.default_rng(500 + s).standard_normal((n, s))
    AP, G, path = gpu_spmm_gram(M, n, s, P)
    want = po.csrmm("S", n, s, 1.0, rp, ci, v, P, 0.0, np.zeros((n, s)))
    assert np.array_equal(AP, want)
    if kind == "random":
        assert path in (0, 1)
    ref = P.T @ want
    assert np.max(np.abs(G - ref)) <= 1e-13 * np.max(np.abs(ref))
