"""The allocation-free CG entry point bkh_cg_noalloc (no counterpart in the
reference: for callers that cannot allocate device memory) against bkh_cg on the
same inputs, through the C-ABI. Both run the same fused kernels with the same
partial-sum layout, so the solutions must agree bit for bit; two calls in a row,
the second on a smaller system, share the static scratch array. Rows beyond the
scratch limit are rejected before any device work."""
import ctypes

import numpy as np
import pytest

from bksp_amd import device as dv
from bksp_amd import gen
from bksp_amd._lib import BkhInfo, load
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def noalloc():
    lib = load()
    return lib.bkh_cg_noalloc, lib.bkh_cg_noalloc_max_rows()


def test_noalloc_cg_equals_bkh_cg_on_two_calls():
    f, max_rows = noalloc()
    assert max_rows == 1 << 20
    for m, seed in ((64, 11), (40, 12)):           # the second call is smaller
        print(f"bkh_cg_noalloc vs bkh_cg: Poisson {m}^3", flush=True)
        n, rp, ci, v = gen.poisson3d(m, lower=True)
        b = gen.normal(seed, n)
        M = dv.Matrix(n, rp, ci, v, symmetric=True)
        db = dv.DeviceArray.from_numpy(b)
        dx = dv.DeviceArray.from_numpy(np.zeros(n))
        ref = dv.cg(M, db, dx, 1e-9, 5000)
        xref = dx.numpy()
        dy = dv.DeviceArray.from_numpy(np.zeros(n))
        info = BkhInfo()
        ret = f(M.handle, db.ptr, dy.ptr, 1e-9, 5000, ctypes.byref(info))
        assert ret == ref.ret and ret >= 0, (ret, ref.ret, dv.last_error())
        assert info.iters == ref.iters and info.status == ref.status
        x = dy.numpy()
        assert np.array_equal(x, xref)
        assert info.true_res < 1.5e-9 and abs(info.true_res - ref.true_res) <= 1e-12 * ref.true_res + 1e-30
        # and the oracle's CG on the same system (x at the device's index)
        cret, _ = po.cg(n, rp, ci, v, b, 1e-9, 5000)
        assert abs(ret - cret) <= max(1, int(0.05 * cret))


def test_noalloc_cg_limits_and_codes():
    f, max_rows = noalloc()
    n, rp, ci, v = gen.poisson3d(8, lower=True)
    M = dv.Matrix(n, rp, ci, v, symmetric=True)
    db = dv.DeviceArray.from_numpy(np.ones(n))
    dx = dv.DeviceArray.from_numpy(np.zeros(n))
    info = BkhInfo()
    assert f(M.handle, db.ptr, dx.ptr, 1e-8, 0, ctypes.byref(info)) == 0          # cg.c:15,67-70
    assert f(M.handle, db.ptr, dx.ptr, 1e-30, 1, ctypes.byref(info)) == -1        # NOT_CONVERGED
    assert f(None, db.ptr, dx.ptr, 1e-8, 10, ctypes.byref(info)) == -6            # BKH_BAD_ARGUMENT
    m = 102                                                                        # 102^3 > 2^20 rows
    n, rp, ci, v = gen.poisson3d(m, lower=True)
    assert n > max_rows
    M = dv.Matrix(n, rp, ci, v, symmetric=True)
    db = dv.DeviceArray.from_numpy(np.ones(n))
    dx = dv.DeviceArray.from_numpy(np.zeros(n))
    assert f(M.handle, db.ptr, dx.ptr, 1e-8, 10, ctypes.byref(info)) == -6
    assert "scratch limit" in dv.last_error()
