"""GPU block CG (dblcg / bkh_blcg) against the oracle's restatement of src/blcg.c.

The device QR is CholeskyQR2 (positive-diagonal R) where the reference uses
Householder; X is invariant under the sign choice, so trajectories agree to
rounding. Both sides use the calloc reading of blcg.c's first direction (P0 = 2V)."""
import numpy as np
import pytest

from bksp_amd import bksp
from bksp_amd import device as dv
from bksp_amd import gen
from conftest import count_index_mismatch, expand_lower, oracle_at_index
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def close_index(a, b):
    return abs(a - b) <= max(1, int(0.05 * max(a, b)))


def gpu_blcg(n, rp, ci, v, B, tol, max_iter, fix_p0=0):
    M = dv.Matrix(n, rp, ci, v, symmetric=True)
    s = B.shape[1]
    dB = dv.DeviceArray.from_numpy(np.ascontiguousarray(B.T))      # column-major m x s
    dX = dv.DeviceArray.from_numpy(np.zeros((s, n)))
    info = dv.blcg(M, s, dB, dX, tol, max_iter, fix_blcg_p0=fix_p0)
    return info, dX.numpy().T


def test_reference_scenario_16_unit_rhs(bfwb398):
    """tests/test.c:132-153: bfwb398, B = [e1..e16], tol 1e-14, max_iter 500."""
    n, rp, ci, v, _ = bfwb398
    B = np.zeros((n, 16)); B[np.arange(16), np.arange(16)] = 1.0
    A = bksp.dspmat(n, rp, ci, v, symmetric=True)
    Bm, Xm = bksp.Mat(B), bksp.Mat(np.zeros((n, 16)))
    ret = bksp.dblcg(A, Bm, 1e-14, 500, Xm)
    cret, cX = po.blcg(n, rp, ci, v, B, 1e-14, 500)
    Af = expand_lower(n, rp, ci, v)
    res = np.linalg.norm(B - Af @ Xm.a) / np.linalg.norm(B)
    assert res < 1e-9, (ret, cret, res)
    assert ret >= 0 and cret >= 0, (ret, cret)
    assert abs(ret - cret) <= max(2, int(0.1 * cret)), (ret, cret)
    # both reach 1e-14: X agrees with the oracle's converged X, and with its
    # iterate at the device's break index
    assert rel(Xm.a, cX) < 1e-8
    count_index_mismatch(ret, cret)
    Xk = oracle_at_index(lambda t, k: po.blcg(n, rp, ci, v, B, t, k), ret)
    assert rel(Xm.a, Xk) < 1e-8


@pytest.mark.parametrize("fix_p0", [0, 1])
def test_first_steps_match_oracle(bfwb398, fix_p0):
    n, rp, ci, v, _ = bfwb398
    B = gen.normal(31, n * 4).reshape(4, n).T.copy()
    for k in (1, 2, 5):
        _, X = gpu_blcg(n, rp, ci, v, B, 1e-300, k, fix_p0)
        _, cX = po.blcg(n, rp, ci, v, B, 1e-300, k, p0_mode=fix_p0)
        assert rel(X, cX) < 1e-9, (k, rel(X, cX))
    # one step with P0 = 2V is half the block Galerkin step (closed form)
    A = expand_lower(n, rp, ci, v).toarray()
    V = np.linalg.qr(B)[0]
    full = V @ np.linalg.solve(V.T @ A @ V, V.T @ B)
    _, X1 = gpu_blcg(n, rp, ci, v, B, 1e-300, 1, fix_p0)
    assert rel(X1, full * (1.0 if fix_p0 else 0.5)) < 1e-9


@pytest.mark.parametrize("m,s", [(24, 8), (40, 8), (20, 3), (20, 16)])
def test_poisson_parity(m, s):
    n, rp, ci, v = gen.poisson3d(m, lower=True)
    B = gen.normal(202, n * s).reshape(s, n).T.copy()
    info, X = gpu_blcg(n, rp, ci, v, B, 1e-8, 3000)
    cret, cX = po.blcg(n, rp, ci, v, B, 1e-8, 3000)
    assert info.ret >= 0 and cret >= 0
    assert close_index(info.ret, cret), (info.ret, cret)
    assert info.true_res < 2e-8
    # the reported FP64 true residual is the host's |B - A X|_F / |B|_F
    Af = expand_lower(n, rp, ci, v)
    res = np.linalg.norm(B - Af @ X) / np.linalg.norm(B)
    assert res > 0 and abs(info.true_res - res) <= 1e-6 * res, (info.true_res, res)
    count_index_mismatch(info.ret, cret)
    Xk = oracle_at_index(lambda t, k: po.blcg(n, rp, ci, v, B, t, k), info.ret)
    assert rel(X, Xk) < 1e-8


def test_cfg2_full_size_blcg_s8():
    """cfg2: 128^3 Poisson, s = 8 RHS N(0,1) seed 202, tol 1e-8, reference P0 = 2V and
    the corrected P0 = V (reported separately, never mixed into the parity claim)."""
    n, rp, ci, v = gen.poisson3d(128, lower=True)
    B = gen.normal(202, n * 8).reshape(8, n).T.copy()
    info, X = gpu_blcg(n, rp, ci, v, B, 1e-8, 5000)
    assert info.ret >= 0 and info.true_res < 2e-8
    info1, _ = gpu_blcg(n, rp, ci, v, B, 1e-8, 5000, fix_p0=1)
    assert info1.ret >= 0 and info1.true_res < 2e-8
    assert info1.ret < info.ret
    print(f"cfg2 block CG s=8: P0=2V {info.ret} iters {info.time_ms:.1f} ms; "
          f"P0=V {info1.ret} iters {info1.time_ms:.1f} ms")


def test_cfg2_full_size_blcg_first_iterates():
    """cfg2's block CG at its own size (128^3 Poisson, s = 8, RHS N(0,1) seed 202,
    P0 = 2V as the reference): the device iterates after 1 and 5 iterations equal the
    oracle's (src/blcg.c:66-106) to rounding."""
    n, rp, ci, v = gen.poisson3d(128, lower=True)
    B = gen.normal(202, n * 8).reshape(8, n).T.copy()
    M = dv.Matrix(n, rp, ci, v, symmetric=True)
    dB = dv.DeviceArray.from_numpy(np.ascontiguousarray(B.T))
    for k in (1, 5):
        dX = dv.DeviceArray.from_numpy(np.zeros((8, n)))
        info = dv.blcg(M, 8, dB, dX, 1e-300, k)
        assert info.ret == -1
        X = dX.numpy().T
        _, cX = po.blcg(n, rp, ci, v, B, 1e-300, k)
        err = rel(X, cX)
        print(f"cfg2 block CG s=8 after {k} iterations: |X - X_oracle|_F / |X_oracle|_F = {err:.2e}")
        assert err < 1e-10, (k, err)


@pytest.mark.parametrize("m,s,fix_p0,path", [(24, 8, 0, 0), (24, 8, 1, 0), (20, 3, 0, 0), (24, 8, 0, 1)])
def test_fp32_ir_small(m, s, fix_p0, path):
    """FP32 block CG with FP64 reliable updates (path 0; path 1: restarted
    refinement) reaches the FP64 tolerance; the solution agrees with the oracle's
    FP64 one to cond(A) * tol."""
    n, rp, ci, v = gen.poisson3d(m, lower=True)
    B = gen.normal(202, n * s).reshape(s, n).T.copy()
    M = dv.Matrix(n, rp, ci, v, symmetric=True, fp32=True)
    dB = dv.DeviceArray.from_numpy(np.ascontiguousarray(B.T))
    dX = dv.DeviceArray.from_numpy(np.zeros((s, n)))
    info = dv.blcg(M, s, dB, dX, 1e-8, 3000, precision=dv.BKH_FP32_IR, fix_blcg_p0=fix_p0, path=path)
    X = dX.numpy().T
    cret, cX = po.blcg(n, rp, ci, v, B, 1e-8, 3000, p0_mode=fix_p0)
    Af = expand_lower(n, rp, ci, v)
    res = np.linalg.norm(B - Af @ X) / np.linalg.norm(B)
    print(f"m={m} s={s} p0={fix_p0} path={path}: fp32-ir inner {info.inner_iters} passes {info.passes} "
          f"(fp64 oracle {cret}) true res {res:.2e}")
    assert info.ret >= 0 and info.passes >= 1
    assert res < 1e-8 and abs(info.true_res - res) < 1e-10
    assert rel(X, cX) < 1e-5


def test_cfg2_full_size_blcg_fp32_ir():
    """cfg2 block CG s = 8 with the FP32 inner solve: true residual below 1e-8."""
    n, rp, ci, v = gen.poisson3d(128, lower=True)
    B = gen.normal(202, n * 8).reshape(8, n).T.copy()
    M = dv.Matrix(n, rp, ci, v, symmetric=True, fp32=True)
    dB = dv.DeviceArray.from_numpy(np.ascontiguousarray(B.T))
    for fix in (0, 1):
        dX = dv.DeviceArray.from_numpy(np.zeros((8, n)))
        info = dv.blcg(M, 8, dB, dX, 1e-8, 5000, precision=dv.BKH_FP32_IR, fix_blcg_p0=fix)
        print(f"cfg2 block CG s=8 fp32-ir P0={'V' if fix else '2V'}: inner {info.inner_iters} "
              f"passes {info.passes} {info.time_ms:.1f} ms true res {info.true_res:.2e}")
        assert info.ret >= 0 and info.true_res < 1e-8


def test_codes():
    n, rp, ci, v = gen.poisson3d(6, lower=True)
    B = np.ones((n, 2))
    A = bksp.dspmat(n, rp, ci, v, symmetric=True)
    assert bksp.dblcg(bksp.dspmat(n, rp, ci, v), bksp.Mat(B), 1e-8, 10, bksp.Mat(np.zeros((n, 2)))) \
        == bksp.INVALID_MATRIX_TYPE
    assert bksp.dblcg(A, bksp.Mat(B), 1e-8, 0, bksp.Mat(np.zeros((n, 2)))) == 0
    assert bksp.dblcg(A, bksp.Mat(np.ones((n, 17))), 1e-8, 5, bksp.Mat(np.zeros((n, 17)))) \
        == bksp.INVALID_PARAMETERS


def test_dropin_fp32ir_environment(monkeypatch):
    """BKSP_PRECISION=fp32ir routes the drop-in dblcg and bksp_dcg through the FP32
    inner / FP64 refinement path; both still meet tol in the FP64 true residual."""
    monkeypatch.setenv("BKSP_PRECISION", "fp32ir")
    n, rp, ci, v = gen.poisson3d(20, lower=True)
    Af = expand_lower(n, rp, ci, v)
    A = bksp.dspmat(n, rp, ci, v, symmetric=True)
    B = gen.normal(7, n * 4).reshape(4, n).T.copy()
    Xm = bksp.Mat(np.zeros((n, 4)))
    assert bksp.dblcg(A, bksp.Mat(B), 1e-9, 2000, Xm) >= 0
    assert np.linalg.norm(B - Af @ Xm.a) / np.linalg.norm(B) < 1e-9
    x = np.zeros(n)
    assert bksp.bksp_dcg(A, B[:, 0].copy(), 1e-9, 2000, x) >= 0
    assert np.linalg.norm(B[:, 0] - Af @ x) / np.linalg.norm(B[:, 0]) < 1e-9


@pytest.mark.parametrize("m,s", [(15, 8), (24, 4), (20, 3)])
def test_blcg_dia_layout_bitwise_equal_to_sell(m, s):
    """The block SpMM on the slice-diagonal copy (s = 4, 8: 16-byte row segments per
    lane) gives the solutions of the SELL-64 SpMM bit for bit, FP64 and FP32-IR; s = 3
    has no such split and runs the SELL kernels on both matrices."""
    import os
    n, rp, ci, v = gen.poisson3d(m, lower=True)
    B = gen.normal(77, n * s).reshape(s, n).T.copy()
    A = dv.Matrix(n, rp, ci, v, symmetric=True, fp32=True)
    os.environ["BKH_DIA"] = "0"
    try:
        C = dv.Matrix(n, rp, ci, v, symmetric=True, fp32=True)
    finally:
        del os.environ["BKH_DIA"]
    assert A.dia_entries > 0 and C.dia_entries == 0
    dB = dv.DeviceArray.from_numpy(np.ascontiguousarray(B.T))
    for prec in (0, dv.BKH_FP32_IR):
        out = []
        for M in (A, C):
            dX = dv.DeviceArray.from_numpy(np.zeros((s, n)))
            info = dv.blcg(M, s, dB, dX, 1e-9, 3000, precision=prec)
            assert info.ret >= 0 and info.true_res < 1e-9
            out.append((info.ret, info.inner_iters, dX.numpy()))
        assert out[0][:2] == out[1][:2]
        assert np.array_equal(out[0][2], out[1][2])


@pytest.mark.parametrize("m,s", [(24, 8), (20, 4), (20, 3), (16, 1)])
def test_small_dense_steps_register_form_bitwise(m, s):
    """The s x s steps of the iteration (inverse of P^T A P, the two CholeskyQR
    factors, their products and ||C||_F; src/blcg.c:78-105) in the register form equal
    the LDS form bit for bit: same iteration counts and the same X, FP64 and FP32-IR."""
    n, rp, ci, v = gen.poisson3d(m, lower=True)
    B = gen.normal(91, n * s).reshape(s, n).T.copy()
    M = dv.Matrix(n, rp, ci, v, symmetric=True, fp32=True)
    dB = dv.DeviceArray.from_numpy(np.ascontiguousarray(B.T))
    lib = dv.load()
    try:
        for prec in (0, dv.BKH_FP32_IR):
            out = []
            for lds in (0, 1):
                lib.bkh_debug_fin_lds_form(lds)
                dX = dv.DeviceArray.from_numpy(np.zeros((s, n)))
                info = dv.blcg(M, s, dB, dX, 1e-9, 3000, precision=prec)
                assert info.ret >= 0 and info.true_res < 1e-9
                out.append((info.ret, info.inner_iters, dX.numpy()))
            assert out[0][:2] == out[1][:2]
            assert np.array_equal(out[0][2], out[1][2])
    finally:
        lib.bkh_debug_fin_lds_form(0)
