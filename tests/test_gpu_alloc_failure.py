"""Device-allocation failure injection. Every device allocation of the library goes
through one wrapper; bkh_debug_fail_alloc(k) makes the k-th next one fail as if HBM
were exhausted. For each workload, k = 1, 2, ... until the workload no longer reaches
a k-th allocation: every run must end in the device-error return (raised here as
BkspDeviceError) or, where a layout is optional, succeed without it -- never crash --
and afterwards the same workload must run normally and give the right answer."""
import numpy as np
import pytest

from bksp_amd import bksp
from bksp_amd import device as dv
from bksp_amd import fsai, gen, load
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def sweep(op, max_k=400):
    """Runs op() with the k-th allocation failing for k = 1, 2, ...; returns (failures
    that fired, runs that still succeeded after a failure)."""
    L = load()
    fired = survived = 0
    for k in range(1, max_k):
        L.bkh_debug_fail_alloc(k)
        try:
            op()
            ok = True
        except dv.BkspDeviceError:
            ok = False
        finally:
            pending = L.bkh_debug_fail_alloc_pending()
            L.bkh_debug_fail_alloc(0)
        if pending > 0:                 # op made fewer than k allocations
            assert ok, k
            return fired, survived
        fired += 1
        survived += ok
    raise AssertionError(f"more than {max_k} allocations")


def test_cg_and_fp32_refinement():
    n, rp, ci, v = gen.poisson3d(12, lower=True)
    b = gen.normal(11, n)

    def op():
        M = dv.Matrix(n, rp, ci, v, symmetric=True, fp32=True)
        db, dx = dv.DeviceArray.from_numpy(b), dv.DeviceArray((n,))
        dx.zero()
        dv.cg(M, db, dx, 1e-8, 500)
        dx.zero()
        dv.cg(M, db, dx, 1e-8, 500, precision=dv.BKH_FP32_IR)
        return dx.numpy()

    fired, survived = sweep(op)
    print(f"cg: {fired} injected failures, {survived} survived by fallback")
    assert fired > 5
    x = op()
    cret, cx = po.cg(n, rp, ci, v, b, 1e-8, 500)
    assert np.linalg.norm(x - cx) / np.linalg.norm(cx) < 1e-5


def test_block_bicgstabl_shifted_and_batched():
    n, rp, ci, v = gen.convdiff3d(8, 0.3, 0.2, 0.1)
    b = gen.normal(12, n)
    B = np.asfortranarray(np.random.default_rng(13).standard_normal((n, 4)))
    systems = [gen.random_rows(900 + q, m) for q, m in enumerate((50, 300, 120))]
    bb = gen.normal(14, sum(s[0] for s in systems))

    def op():
        M = dv.Matrix(n, rp, ci, v, fp32=True)
        db, dx = dv.DeviceArray.from_numpy(b), dv.DeviceArray((n,))
        dx.zero()
        dv.bicgstabl(M, db, 2, dx, 1e-8, 300)
        dB, dX = dv.DeviceArray.from_numpy(B.ravel(order="F")), dv.DeviceArray((n * 4,))
        dX.zero()
        dv.blbicgstabl(M, 4, dB, 2, dX, 1e-8, 300)
        dxs = dv.DeviceArray((n * 3,))
        dxs.zero()
        dv.shbicgstabl(M, db, 2, [0.1, 0.5], dxs, 1e-8, 300, precision=dv.BKH_FP32_IR)
        Bt = dv.Batch(systems, fp32=True)
        dbb, dxb = dv.DeviceArray.from_numpy(bb), dv.DeviceArray((Bt.N,))
        dxb.zero()
        Bt.bicgstabl(dbb, 2, dxb, 1e-8, 300, precision=dv.BKH_FP32_IR)
        dxb.zero()
        Bt.bicgstabl(dbb, 2, dxb, 1e-8, 300, path=2)
        return dx.numpy()

    fired, survived = sweep(op)
    print(f"general solvers: {fired} injected failures, {survived} survived by fallback")
    assert fired > 10
    x = op()
    cret, cx = po.bicgstabl(n, rp, ci, v, b, 2, 1e-8, 300)
    assert np.linalg.norm(x - cx) / np.linalg.norm(cx) < 1e-6


def test_block_cg_and_fsai():
    n, rp, ci, v = gen.poisson3d(10, lower=True)
    B = np.asfortranarray(np.random.default_rng(15).standard_normal((n, 8)))
    fn, low = fsai.field_system(6100, 12)
    fb = gen.normal(16, fn)

    def op():
        M = dv.Matrix(n, rp, ci, v, symmetric=True, fp32=True)
        dB, dX = dv.DeviceArray.from_numpy(B.ravel(order="F")), dv.DeviceArray((n * 8,))
        dX.zero()
        dv.blcg(M, 8, dB, dX, 1e-8, 300)
        dX.zero()
        dv.blcg(M, 8, dB, dX, 1e-8, 300, precision=dv.BKH_FP32_IR)
        F = dv.FsaiField(fn, *low)
        model = dv.FsaiModel(fsai.init_params())
        F.factor(model)
        d_b, d_x = dv.DeviceArray.from_numpy(fb), dv.DeviceArray((fn,))
        d_x.zero()
        info = F.pcg(d_b, d_x, 1e-8, 2000)
        return info

    fired, survived = sweep(op)
    print(f"block CG / FSAI: {fired} injected failures, {survived} survived by fallback")
    assert fired > 10
    info = op()
    assert info.ret >= 0 and info.true_res < 1e-8


def test_dropin_entry_points(monkeypatch):
    monkeypatch.setenv("BKSP_CACHE", "0")          # every call uploads (and allocates)
    n, rp, ci, v = gen.poisson3d(10, lower=True)
    b = gen.normal(17, n)
    A = bksp.dspmat(n, rp, ci, v, symmetric=True)
    L = load()
    rets = []
    for k in range(1, 60):
        x = np.zeros(n)
        L.bkh_debug_fail_alloc(k)
        try:
            ret = bksp.bksp_dcg(A, b, 1e-8, 500, x)
        finally:
            pending = L.bkh_debug_fail_alloc_pending()
            L.bkh_debug_fail_alloc(0)
        if pending > 0:
            assert ret >= 0
            break
        assert ret == -5 or ret >= 0, (k, ret)       # the device-error code, or a fallback
        rets.append(ret)
    assert -5 in rets
    x = np.zeros(n)
    assert bksp.bksp_dcg(A, b, 1e-8, 500, x) >= 0
    cret, cx = po.cg(n, rp, ci, v, b, 1e-8, 500)
    assert np.linalg.norm(x - cx) / np.linalg.norm(cx) < 1e-8
