"""GPU BiCGSTAB(l) (dbicgstabl / bkh_bicgstabl) against the oracle's restatement of
src/bicgstabl.c. The device forms the MR step from a fused Gram matrix (same
algebra as the reference's modified Gram-Schmidt), so parity is: outer index
within +-5% (at least +-1), x within 1e-8 relative of the oracle's iterate at the
device's break index, and a true FP64 residual at the tolerance."""
import numpy as np
import pytest

from bksp_amd import bksp
from bksp_amd import device as dv
from bksp_amd import gen
from conftest import count_index_mismatch, oracle_at_index, to_scipy
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def close_index(a, b):
    return abs(a - b) <= max(1, int(0.05 * max(a, b)))


def close_index_fixture(a, b):
    """The reference's fixture scenarios run at tol 1e-14, where the last digits of
    the recursive residual decide the stop: two algebraically identical reference
    codes (bicgstabl.c and shbicgstabl.c's seed) already differ by up to 3 outer
    iterations there on pde2961, so these scenarios allow +-max(2, 10%)."""
    return abs(a - b) <= max(2, int(0.1 * max(a, b)))


def cfg1():
    """64x64 convection-diffusion, velocity (10, 20), h = 1/65; x_true ~ U[-1,1] seed 101."""
    n, rp, ci, v = gen.convdiff2d(64, 10.0, 20.0)
    assert n == 4096 and rp[-1] == 20224
    xt = gen.uniform(101, n, -1.0, 1.0)
    b = po.csrgemv("n", n, rp, ci, v, xt)
    return n, rp, ci, v, xt, b


def test_cfg1_dropin_parity():
    n, rp, ci, v, xt, b = cfg1()
    A = bksp.dspmat(n, rp, ci, v)
    x = np.zeros(n)
    ret = bksp.dbicgstabl(A, b, 2, 1e-10, 1000, x)
    cret, cx = po.bicgstabl(n, rp, ci, v, b, 2, 1e-10, 1000)
    assert ret >= 0 and cret >= 0
    assert close_index(ret, cret), (ret, cret)
    S = to_scipy(n, rp, ci, v)
    assert np.linalg.norm(b - S @ x) / np.linalg.norm(b) < 2e-10
    count_index_mismatch(ret, cret)
    xk = oracle_at_index(lambda t, k: po.bicgstabl(n, rp, ci, v, b, 2, t, k), ret)
    assert rel(x, xk) < 1e-8
    assert rel(x, xt) < 1e-6


def test_reference_scenario_l4(pde2961):
    """tests/test.c:193-209: pde2961, b = e1, l = 4, tol 1e-14, max_iter 500."""
    n, rp, ci, v, _ = pde2961
    b = np.zeros(n); b[0] = 1.0
    A = bksp.dspmat(n, rp, ci, v)
    x = np.zeros(n)
    ret = bksp.dbicgstabl(A, b, 4, 1e-14, 500, x)
    cret, cx = po.bicgstabl(n, rp, ci, v, b, 4, 1e-14, 500)
    S = to_scipy(n, rp, ci, v)
    assert (ret >= 0) == (cret >= 0)
    if ret >= 0:
        assert close_index_fixture(ret, cret), (ret, cret)
    assert np.linalg.norm(b - S @ x) < 1e-9
    assert rel(x, cx) < 1e-7
    k = ret if ret >= 0 else 499
    xk = oracle_at_index(lambda t, m: po.bicgstabl(n, rp, ci, v, b, 4, t, m), k)
    assert rel(x, xk) < 1e-8, rel(x, xk)


@pytest.mark.parametrize("l", [1, 2, 3, 4, 6, 8])
def test_degrees(pde2961, l):
    n, rp, ci, v, _ = pde2961
    b = gen.normal(7, n)
    M = dv.Matrix(n, rp, ci, v)
    db, dx = dv.DeviceArray.from_numpy(b), dv.DeviceArray.from_numpy(np.zeros(n))
    info = dv.bicgstabl(M, db, l, dx, 1e-10, 2000)
    cret, cx = po.bicgstabl(n, rp, ci, v, b, l, 1e-10, 2000)
    assert info.ret >= 0 and cret >= 0
    assert abs(info.ret - cret) <= max(2, int(0.1 * cret)), (info.ret, cret)
    assert info.true_res < 5e-10
    # both stop below 1e-10 relative residual, so their x agree to ~cond * 1e-10
    assert rel(dx.numpy(), cx) < 1e-6


def test_first_outer_iteration_matches_oracle(pde2961):
    """With max_iter = k the device and oracle iterates agree to rounding."""
    n, rp, ci, v, _ = pde2961
    b = gen.normal(8, n)
    M = dv.Matrix(n, rp, ci, v)
    for l in (2, 4):
        for k in (1, 2, 3):
            db, dx = dv.DeviceArray.from_numpy(b), dv.DeviceArray.from_numpy(np.zeros(n))
            dv.bicgstabl(M, db, l, dx, 1e-300, k)
            _, cx = po.bicgstabl(n, rp, ci, v, b, l, 1e-300, k)
            assert rel(dx.numpy(), cx) < 1e-10, (l, k)


def test_identity_exact_and_edge_cases():
    n = 1000
    rp = np.arange(n + 1, dtype=np.int32)
    ci = np.arange(n, dtype=np.int32)
    A = bksp.dspmat(n, rp, ci, np.ones(n))
    b = gen.normal(9, n)
    x = np.zeros(n)
    # r becomes exactly 0 in the first BiCG step: the device stops cleanly at index 0
    # (the reference divides 0/0 in the MR step and returns max_iter with a NaN x)
    assert bksp.dbicgstabl(A, b, 1, 1e-10, 50, x) == 0
    assert np.array_equal(x, b)
    assert bksp.dbicgstabl(A, b, 2, 1e-10, 0, np.zeros(n)) == 0
    n2, rp2, ci2, v2 = gen.convdiff2d(16, 10.0, 20.0)
    A2 = bksp.dspmat(n2, rp2, ci2, v2)
    assert bksp.dbicgstabl(A2, np.ones(n2), 2, 1e-30, 1, np.zeros(n2)) == bksp.NOT_CONVERGED
    bad = np.ones(n2); bad[3] = np.nan
    assert bksp.dbicgstabl(A2, bad, 2, 1e-10, 20, np.zeros(n2)) == bksp.DIVERGENCE


def test_cfg1_fp32_ir():
    n, rp, ci, v, xt, b = cfg1()
    M = dv.Matrix(n, rp, ci, v, fp32=True)
    db, dx = dv.DeviceArray.from_numpy(b), dv.DeviceArray.from_numpy(np.zeros(n))
    info = dv.bicgstabl(M, db, 2, dx, 1e-10, 2000, precision=dv.BKH_FP32_IR)
    assert info.ret >= 0 and info.true_res < 1e-10
    assert rel(dx.numpy(), xt) < 1e-6


@pytest.mark.parametrize("path", [1, 2])
def test_cfg1_both_paths(path):
    """cfg1 through the multi-kernel graph path (1) and the single persistent
    workgroup path (2): both meet the oracle's index and the true residual."""
    n, rp, ci, v, xt, b = cfg1()
    M = dv.Matrix(n, rp, ci, v)
    db, dx = dv.DeviceArray.from_numpy(b), dv.DeviceArray.from_numpy(np.zeros(n))
    info = dv.bicgstabl(M, db, 2, dx, 1e-10, 1000, path=path)
    cret, cx = po.bicgstabl(n, rp, ci, v, b, 2, 1e-10, 1000)
    assert info.ret >= 0 and close_index(info.ret, cret), (path, info.ret, cret)
    assert info.true_res < 2e-10
    assert rel(dx.numpy(), xt) < 1e-6
    xk = oracle_at_index(lambda t, k: po.bicgstabl(n, rp, ci, v, b, 2, t, k), info.ret)
    assert rel(dx.numpy(), xk) < 1e-8, (path, rel(dx.numpy(), xk))
    for k in (1, 4):
        dx.upload(np.zeros(n))
        dv.bicgstabl(M, db, 2, dx, 1e-300, k, path=path)
        _, ck = po.bicgstabl(n, rp, ci, v, b, 2, 1e-300, k)
        assert rel(dx.numpy(), ck) < 1e-10, (path, k)


@pytest.mark.parametrize("l", [1, 2])
def test_register_kernel_matches_global_memory_kernel(monkeypatch, l):
    """The single-workgroup path keeps the Krylov vectors in registers for these
    sizes; BKH_SMALL_GLOBAL=1 selects its global-memory twin. Same iterates (the
    dot products are summed in a different order: 1e-12), same convergence index."""
    n, rp, ci, v, xt, b = cfg1()
    M = dv.Matrix(n, rp, ci, v)
    db = dv.DeviceArray.from_numpy(b)
    res = {}
    for mode in ("reg", "global"):
        if mode == "global":
            monkeypatch.setenv("BKH_SMALL_GLOBAL", "1")
        xs = []
        for k in (1, 5):
            dx = dv.DeviceArray.from_numpy(np.zeros(n))
            dv.bicgstabl(M, db, l, dx, 1e-300, k, path=2)
            xs.append(dx.numpy())
        dx = dv.DeviceArray.from_numpy(np.zeros(n))
        info = dv.bicgstabl(M, db, l, dx, 1e-10, 2000, path=2)
        res[mode] = (xs, info)
    (xr, ir), (xg, ig) = res["reg"], res["global"]
    for a, c in zip(xr, xg):
        assert rel(a, c) < 1e-12
    assert ir.ret >= 0 and abs(ir.ret - ig.ret) <= 1 and ir.true_res < 2e-10
    print(f"l={l}: register kernel {ir.ret} outer, {ir.time_ms:.3f} ms; global-memory kernel {ig.ret}, "
          f"{ig.time_ms:.3f} ms")


@pytest.mark.parametrize("l", [1, 2])
def test_register_matrix_kernel_bitwise(monkeypatch, l):
    """On the slice-diagonal copy the single-workgroup kernel also holds the matrix
    values in registers (BKH_SMALL_NODIA=1 selects the kernel that streams them from
    L2 each SpMV): the same sums in the same order, so bitwise the same iterates and
    the same convergence index."""
    n, rp, ci, v, xt, b = cfg1()
    M = dv.Matrix(n, rp, ci, v)
    assert M.dia_entries > 0
    db = dv.DeviceArray.from_numpy(b)
    res = {}
    for mode in ("dia", "stream"):
        if mode == "stream":
            monkeypatch.setenv("BKH_SMALL_NODIA", "1")
        out = []
        for k in (1, 5):
            dx = dv.DeviceArray.from_numpy(np.zeros(n))
            dv.bicgstabl(M, db, l, dx, 1e-300, k, path=2)
            out.append(dx.numpy())
        dx = dv.DeviceArray.from_numpy(np.zeros(n))
        info = dv.bicgstabl(M, db, l, dx, 1e-10, 2000, path=2)
        out.append(dx.numpy())
        res[mode] = (out, info)
    (xd, idia), (xs, ist) = res["dia"], res["stream"]
    for a, c in zip(xd, xs):
        assert np.array_equal(a, c)
    assert idia.ret == ist.ret and idia.ret >= 0 and idia.true_res < 2e-10
    print(f"l={l}: matrix in registers {idia.ret} outer, {idia.time_ms:.3f} ms; streamed {ist.time_ms:.3f} ms")
