"""The slice-diagonal copy of banded matrices against the SELL-64 copy: every solver
that runs on it gives the same iterates bit for bit (a value-0 position adds a 0
product at a clamped column, which leaves a finite row sum unchanged, and the
positions keep each row's ascending order), and the public SpMV entry points keep the
SELL-64 kernel, so a non-finite x reaches only the rows that reference it, as in
src/matvec.c."""
import os

import numpy as np
import pytest

from bksp_amd import device as dv
from bksp_amd import gen

pytestmark = pytest.mark.gpu


def _pair(*args, **kw):
    """The matrix with and without its slice-diagonal copy (BKH_DIA=0 at creation)."""
    A = dv.Matrix(*args, **kw)
    os.environ["BKH_DIA"] = "0"
    try:
        B = dv.Matrix(*args, **kw)
    finally:
        del os.environ["BKH_DIA"]
    assert A.dia_entries > 0 and B.dia_entries == 0
    return A, B


def _same(run, A, B, n):
    out = []
    for M in (A, B):
        dx = dv.DeviceArray.from_numpy(np.zeros(n))
        info = run(M, dx)
        out.append((info.ret, info.inner_iters, info.true_res, dx.numpy()))
    assert out[0][:3] == out[1][:3], (out[0][:3], out[1][:3])
    assert np.array_equal(out[0][3], out[1][3])
    return out[0]


@pytest.mark.parametrize("l,prec", [(2, 0), (4, 0), (2, 1)])
def test_bicgstabl_dia_bitwise(l, prec):
    px, py, pz = gen.cfg3_peclet()
    n, rp, ci, v = gen.convdiff3d(17, px, py, pz)
    A, B = _pair(n, rp, ci, v, fp32=True)
    db = dv.DeviceArray.from_numpy(gen.normal(5, n))
    ret, _, tr, _ = _same(lambda M, dx: dv.bicgstabl(M, db, l, dx, 1e-9, 2000, precision=prec), A, B, n)
    assert ret >= 0 and tr < 1e-9


def test_cr_and_bicg_dia_bitwise():
    px, py, pz = gen.cfg3_peclet()
    n, rp, ci, v = gen.convdiff3d(15, px, py, pz)
    A, B = _pair(n, rp, ci, v, transpose=True)
    db = dv.DeviceArray.from_numpy(gen.normal(6, n))
    ret, _, tr, _ = _same(lambda M, dx: dv.bicg(M, db, dx, 1e-9, 2000), A, B, n)
    assert ret >= 0 and tr < 1e-9
    n, rp, ci, v = gen.poisson3d(15, lower=True)
    A, B = _pair(n, rp, ci, v, symmetric=True)
    db = dv.DeviceArray.from_numpy(gen.normal(7, n))
    ret, _, tr, _ = _same(lambda M, dx: dv.cr(M, db, dx, 1e-9, 2000), A, B, n)
    assert ret >= 0 and tr < 1e-9


@pytest.mark.parametrize("kind", [0, 1])
def test_pcg_dia_bitwise(kind):
    m = 16
    kap = gen.grf_kappa(11, m)
    n, rp, ci, v = gen.fv_diffusion(m, kap, lower=True)
    A, B = _pair(n, rp, ci, v, symmetric=True, fp32=True)
    db = dv.DeviceArray.from_numpy(gen.normal(8, n))
    for path in (1, 2):
        ret, _, tr, _ = _same(lambda M, dx: dv.pcg(M, kind, db, dx, 1e-9, 20000, path=path), A, B, n)
        assert ret >= 0 and tr < 1e-9


@pytest.mark.parametrize("prec", [0, 1])
def test_shifted_dia_bitwise(prec):
    px, py, pz = gen.cfg3_peclet()
    n, rp, ci, v = gen.convdiff3d(14, px, py, pz)
    A, B = _pair(n, rp, ci, v, fp32=True)
    db = dv.DeviceArray.from_numpy(gen.normal(9, n))
    sig = 10.0 ** (-3 + 3 * np.arange(8) / 7.0)
    out = []
    for M in (A, B):
        dx = dv.DeviceArray((n * 9,))
        dx.zero()
        info, tr = dv.shbicgstabl(M, db, 4, sig, dx, 1e-8, 1000, precision=prec)
        out.append((info.ret, info.inner_iters, dx.numpy(), np.asarray(tr)))
    assert out[0][:2] == out[1][:2]
    assert np.array_equal(out[0][2], out[1][2]) and np.array_equal(out[0][3], out[1][3])
    # FP32-IR refines every shift below tol; the FP64 solve stops on the seed's residual
    assert (out[0][3].max() if prec else out[0][3][0]) < 1e-8


def test_public_spmv_keeps_sell_semantics():
    """bkh_spmv on a matrix with a slice-diagonal copy: an infinite x_j gives
    non-finite values only in the rows whose stored entries reference column j."""
    n, rp, ci, v = gen.poisson3d(12, lower=False)
    A = dv.Matrix(n, rp, ci, v)
    assert A.dia_entries > 0
    x = np.ones(n)
    j = 5 * 144 + 3 * 12 + 7
    x[j] = np.inf
    dx, dy = dv.DeviceArray.from_numpy(x), dv.DeviceArray((n,))
    A.spmv(dx, dy)
    y = dy.numpy()
    refs = {i for i in range(n) if j in set(ci[rp[i]:rp[i + 1]].tolist())}
    bad = set(np.nonzero(~np.isfinite(y))[0].tolist())
    assert bad == refs


@pytest.mark.parametrize("path", [2, 1])
def test_fsai_pcg_dia_bitwise(path):
    """Learned-FSAI PCG on a field whose matrix has the slice-diagonal copy against one
    built without it (the factor G keeps the SELL-64 / CSR kernels either way)."""
    from bksp_amd import fsai
    m = 16
    n, (rpl, cil, vl) = fsai.field_system(29, m)
    F1 = dv.FsaiField(n, rpl, cil, vl)
    os.environ["BKH_DIA"] = "0"
    try:
        F2 = dv.FsaiField(n, rpl, cil, vl)
    finally:
        del os.environ["BKH_DIA"]
    M = dv.FsaiModel(fsai.init_params(3) * 2)
    db = dv.DeviceArray.from_numpy(gen.normal(12, n))
    out = []
    for F in (F1, F2):
        F.factor(M)
        dx = dv.DeviceArray.from_numpy(np.zeros(n))
        info = F.pcg(db, dx, 1e-9, 100000, path=path)
        out.append((info.ret, info.true_res, dx.numpy()))
    assert out[0][:2] == out[1][:2]
    assert np.array_equal(out[0][2], out[1][2])
    assert out[0][0] >= 0 and out[0][1] < 1.5e-9


@pytest.mark.parametrize("m", [8, 12])
def test_small_matrices_dia_bitwise(m):
    """n = 512, 1728: slices straddling several grid lines, every offset of the rows
    near both ends clamped. CG (FP64 and FP32-IR) on the copy matches the SELL-64
    kernels bit for bit."""
    n, rp, ci, v = gen.poisson3d(m, lower=True)
    A, B = _pair(n, rp, ci, v, symmetric=True, fp32=True)
    db = dv.DeviceArray.from_numpy(gen.normal(40 + m, n))
    for prec in (0, dv.BKH_FP32_IR):
        ret, _, tr, _ = _same(lambda M, dx: dv.cg(M, db, dx, 1e-10, 500, precision=prec), A, B, n)
        assert ret >= 0 and tr < 1e-10


def test_tiny_matrix_has_no_copy():
    """n = 8: SELL-64 padding (one 64-row slice) exceeds its bound, so there is neither a
    SELL nor a slice-diagonal copy; the solve runs on the CSR kernels."""
    n, rp, ci, v = gen.poisson3d(2, lower=True)
    M = dv.Matrix(n, rp, ci, v, symmetric=True, fp32=True)
    assert M.dia_entries == 0
    db = dv.DeviceArray.from_numpy(gen.normal(41, n))
    dx = dv.DeviceArray.from_numpy(np.zeros(n))
    info = dv.cg(M, db, dx, 1e-12, 100)
    assert info.ret >= 0 and info.true_res < 1e-12


def test_complex_solvers_dia_bitwise():
    """The complex solvers' SpMV on the slice-diagonal copies of A and A^T (a value-0
    position adds (0 + 0i) x) gives the iterates of the SELL-64 kernels bit for bit:
    zbicg (A and A^H), zbicgstabl, zshbicgstabl; the public complex SpMV keeps SELL-64."""
    from conftest import zsystem
    n, rp, ci, v = zsystem(40, seed=6)
    A = dv.ZMatrix(n, rp, ci, v)
    os.environ["BKH_DIA"] = "0"
    try:
        B = dv.ZMatrix(n, rp, ci, v)
    finally:
        del os.environ["BKH_DIA"]
    assert A.dia_entries > 0 and B.dia_entries == 0
    rng = np.random.default_rng(3)
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    db = dv.DeviceArray.from_numpy(b)
    sig = np.array([0.1 + 0.2j, 0.5, 1.0 - 0.3j])

    def run(M, which):
        if which == "shifted":
            dx = dv.DeviceArray.from_numpy(np.zeros(n * 4, np.complex128))
            info, tr = dv.zshbicgstabl(M, db, 2, sig, dx, 1e-10, 1000)
            return info.ret, dx.numpy(), tr
        dx = dv.DeviceArray.from_numpy(np.zeros(n, np.complex128))
        info = dv.zbicg(M, db, dx, 1e-10, 2000) if which == "bicg" else dv.zbicgstabl(M, db, 4, dx, 1e-10, 2000)
        return info.ret, dx.numpy(), np.array([info.true_res])

    for which in ("bicg", "bicgstabl", "shifted"):
        ra, xa, ta = run(A, which)
        rb, xb, tb = run(B, which)
        assert ra == rb and ra >= 0, (which, ra, rb)
        assert np.array_equal(xa, xb) and np.array_equal(ta, tb), which
