"""Batched BiCGSTAB(l) over independent uneven systems (bkh_batch_*), against the
oracle's single-system restatement of src/bicgstabl.c run on each system."""
import numpy as np
import pytest

from bksp_amd import device as dv
from bksp_amd import gen
from conftest import count_index_mismatch, oracle_at_index
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

SIZES = [40, 700, 64, 2500, 1031, 300, 97]


def small_batch(seed=4040):
    return [gen.random_rows(seed + q, n) for q, n in enumerate(SIZES)]


def true_res(sy, b, x):
    n, rp, ci, v = sy
    y = po.csrgemv("n", n, rp, ci, v, x)
    return np.linalg.norm(b - y) / np.linalg.norm(b)


def spmv_both_kernels(B, systems, x):
    """y from the column-chunked and the CSR-stream kernel, each bitwise the oracle's
    row-ordered product."""
    dx, dy = dv.DeviceArray.from_numpy(x), dv.DeviceArray.from_numpy(np.zeros(B.N))
    ys = {}
    for k in (1, 0):
        B.spmv_kernel = k
        dy.upload(np.zeros(B.N))
        B.spmv(dx, dy)
        ys[k] = dy.numpy()
    for q, (n, rp, ci, v) in enumerate(systems):
        o = B.offsets[q]
        ref = po.csrgemv("n", n, rp, ci, v, x[o:o + n])
        assert np.array_equal(ys[1][o:o + n], ref), q
        assert np.array_equal(ys[0][o:o + n], ref), q
    return ys[1], dx, dy


def test_batched_spmv_bitwise_and_tuner():
    systems = small_batch()
    B = dv.Batch(systems, fp32=True)
    assert B.spmv_kernel == 1                      # rows ascend: the chunked copy is the default
    x = gen.normal(7, B.N)
    y, dx, dy = spmv_both_kernels(B, systems, x)
    ms = B.tune(fp32=False)
    assert ms > 0
    layout = B.layout()
    dy.upload(np.zeros(B.N))
    B.spmv(dx, dy)
    assert np.array_equal(dy.numpy(), y), (layout, B.spmv_kernel)   # row order does not depend on blocking
    B.tune(fp32=True)
    print("tuned layout", B.layout(), "chunked" if B.spmv_kernel else "csr-stream")


def test_batched_spmv_chunked_many_chunks_and_items():
    """Systems of several 8192-column chunks and several 4096-row items, beside tiny
    ones; FP64 and FP32 products bitwise against the oracle / the CSR-stream kernel."""
    sizes = [20000, 9000, 8192, 4097, 33, 16385, 100]
    systems = [gen.random_rows(7070 + q, n) for q, n in enumerate(sizes)]
    B = dv.Batch(systems, fp32=True)
    assert B.spmv_kernel == 1
    x = gen.normal(8, B.N)
    spmv_both_kernels(B, systems, x)
    x32 = x.astype(np.float32)
    dx, dy = dv.DeviceArray.from_numpy(x32), dv.DeviceArray.from_numpy(np.zeros(B.N, np.float32))
    ys = {}
    for k in (1, 0):
        B.spmv_kernel = k
        dy.upload(np.zeros(B.N, np.float32))
        B.spmv(dx, dy, fp32=True)
        ys[k] = dy.numpy()
    assert np.array_equal(ys[1], ys[0])
    assert np.all(np.isfinite(ys[1]))


def test_batched_spmv_unsorted_row_falls_back():
    """A row whose columns descend leaves the batch without the chunked copy: the
    CSR-stream kernel runs, and asking for the chunked one is refused."""
    systems = small_batch(8080)
    n, rp, ci, v = systems[1]
    ci, v = ci.copy(), v.copy()
    a, e = rp[5], rp[6]
    assert e - a >= 2
    ci[a:e], v[a:e] = ci[a:e][::-1].copy(), v[a:e][::-1].copy()
    systems[1] = (n, rp, ci, v)
    B = dv.Batch(systems)
    assert B.spmv_kernel == 0
    with pytest.raises(ValueError):
        B.spmv_kernel = 1
    x = gen.normal(9, B.N)
    dx, dy = dv.DeviceArray.from_numpy(x), dv.DeviceArray.from_numpy(np.zeros(B.N))
    B.spmv(dx, dy)
    y = dy.numpy()
    for q, (n, rp, ci, v) in enumerate(systems):
        o = B.offsets[q]
        assert np.array_equal(y[o:o + n], po.csrgemv("n", n, rp, ci, v, x[o:o + n])), q


def test_batched_fp64_matches_oracle_per_system():
    systems = small_batch()
    B = dv.Batch(systems)
    b = gen.normal(405, B.N)
    db, dx = dv.DeviceArray.from_numpy(b), dv.DeviceArray.from_numpy(np.zeros(B.N))
    tol = 1e-10
    info, ret, tr = B.bicgstabl(db, 2, dx, tol, 500)
    x = dx.numpy()
    for q, sy in enumerate(systems):
        n = sy[0]
        o = B.offsets[q]
        cret, cx = po.bicgstabl(n, sy[1], sy[2], sy[3], b[o:o + n], 2, tol, 500)
        assert ret[q] >= 0 and cret >= 0, (q, ret[q], cret)
        assert abs(ret[q] - cret) <= max(1, int(0.05 * cret)), (q, ret[q], cret)
        assert tr[q] < 1.5 * tol
        assert abs(true_res(sy, b[o:o + n], x[o:o + n]) - tr[q]) <= 1e-3 * tr[q] + 1e-18
        count_index_mismatch(ret[q], cret)
        xk = oracle_at_index(lambda t, k: po.bicgstabl(n, sy[1], sy[2], sy[3], b[o:o + n], 2, t, k),
                             int(ret[q]))
        assert np.linalg.norm(x[o:o + n] - xk) / np.linalg.norm(xk) < 1e-8, q
    assert info.ret == ret.max()


@pytest.mark.parametrize("precision", [dv.BKH_FP64, dv.BKH_FP32_IR])
@pytest.mark.parametrize("chunked", [1, 0])
def test_batched_persistent_equals_per_phase_graphs(precision, chunked):
    """The persistent launch runs the per-phase kernels' arithmetic in the same order
    (the same SpMV items or units, tiles and partial rows; with the chunked SpMV its
    workgroups take four 256-thread tiles or systems at a time), so x is bitwise the
    graph path's with the same SpMV kernel. The two SpMV kernels give the same products
    but group the dot-product partials by item or by unit, so across kernels x agrees
    to rounding."""
    systems = small_batch(5150)
    B = dv.Batch(systems, fp32=True)
    b = gen.normal(406, B.N)
    db = dv.DeviceArray.from_numpy(b)
    B.spmv_kernel = 1 - chunked
    dx = dv.DeviceArray.from_numpy(np.zeros(B.N))
    info, ret_o, tr_o = B.bicgstabl(db, 2, dx, 1e-9, 500, precision=precision, path=1)
    x_other = dx.numpy()
    assert (ret_o >= 0).all() and tr_o.max() < 1.5e-9
    B.spmv_kernel = chunked
    out = {}
    for path in (1, 2, 0):
        dx = dv.DeviceArray.from_numpy(np.zeros(B.N))
        info, ret, tr = B.bicgstabl(db, 2, dx, 1e-9, 500, precision=precision, path=path)
        assert dv.Batch.last_path() == (2 if path == 2 else 1), path
        out[path] = (dx.numpy(), ret.copy(), tr.copy())
    for path in (2, 0):
        assert np.array_equal(out[path][0], out[1][0]), path
        assert np.array_equal(out[path][1], out[1][1]), path
    assert (out[2][1] >= 0).all() and out[2][2].max() < 1.5e-9
    assert np.linalg.norm(x_other - out[1][0]) / np.linalg.norm(out[1][0]) < 1e-6


def test_persistent_barrier_wait_bound_aborts_cleanly(monkeypatch):
    """A zero poll bound (set from the host) makes every workgroup that has to wait at a
    grid barrier give up at once: the launch ends, the call reports a device error, and
    the next solve on the same device runs normally."""
    systems = small_batch(6160)
    B = dv.Batch(systems)
    b = gen.normal(407, B.N)
    db = dv.DeviceArray.from_numpy(b)
    dx = dv.DeviceArray.from_numpy(np.zeros(B.N))
    monkeypatch.setenv("BKH_BT_SPIN", "0")
    with pytest.raises(dv.BkspDeviceError, match="grid barrier"):
        B.bicgstabl(db, 2, dx, 1e-9, 500, path=2)
    monkeypatch.delenv("BKH_BT_SPIN")
    dx.upload(np.zeros(B.N))
    info, ret, tr = B.bicgstabl(db, 2, dx, 1e-9, 500, path=2)
    assert dv.Batch.last_path() == 2
    assert (ret >= 0).all() and tr.max() < 1.5e-9


def test_batched_fp32_refined_reaches_fp64_tolerance():
    systems = small_batch()
    B = dv.Batch(systems, fp32=True)
    b = gen.normal(405, B.N)
    db = dv.DeviceArray.from_numpy(b)
    dx64 = dv.DeviceArray.from_numpy(np.zeros(B.N))
    B.bicgstabl(db, 2, dx64, 1e-8, 500)
    dx = dv.DeviceArray.from_numpy(np.zeros(B.N))
    info, ret, tr = B.bicgstabl(db, 2, dx, 1e-8, 500, precision=dv.BKH_FP32_IR)
    assert (ret >= 0).all(), ret
    assert tr.max() < 1e-8
    x, x64 = dx.numpy(), dx64.numpy()
    for q, sy in enumerate(systems):
        o, n = B.offsets[q], sy[0]
        assert true_res(sy, b[o:o + n], x[o:o + n]) < 1e-8
        # both solves meet 1e-8; they agree to about cond * tol (reported)
        assert np.linalg.norm(x[o:o + n] - x64[o:o + n]) / np.linalg.norm(x64[o:o + n]) < 1e-5
    print(f"fp32ir passes={info.passes} inner max={info.inner_iters} {info.time_ms:.2f} ms")


def test_cfg4_full_batch():
    """512 uneven systems, n log-uniform in [2^10, 2^17], BiCGSTAB(2), tol 1e-8."""
    systems = gen.cfg4_batch()
    B = dv.Batch(systems, fp32=True)
    del systems
    B.tune(fp32=True)
    b = gen.normal(405, B.N)
    db = dv.DeviceArray.from_numpy(b)
    dx = dv.DeviceArray.from_numpy(np.zeros(B.N))
    info, ret, tr = B.bicgstabl(db, 2, dx, 1e-8, 1000, precision=dv.BKH_FP32_IR)
    print(f"cfg4 N={B.N} nnz={B.nnz} fp32ir {info.time_ms:.1f} ms passes={info.passes} "
          f"max inner={info.inner_iters} max true res={tr.max():.2e} layout={B.layout()}")
    assert (ret >= 0).all()
    assert tr.max() < 1e-8
    kernel = B.spmv_kernel
    print(f"cfg4 tuned SpMV: {'chunked' if kernel else 'csr-stream'}")
    x_tuned = dx.numpy()
    dx.upload(np.zeros(B.N))
    infop, retp, _ = B.bicgstabl(db, 2, dx, 1e-8, 1000, precision=dv.BKH_FP32_IR, path=2)
    assert dv.Batch.last_path() == 2
    print(f"cfg4 fp32ir persistent launch, tuned SpMV: {infop.time_ms:.1f} ms")
    assert np.array_equal(dx.numpy(), x_tuned) and np.array_equal(retp, ret)
    # the persistent launch against the graph path with the CSR-stream SpMV
    B.spmv_kernel = 0
    dx.upload(np.zeros(B.N))
    info0, ret, _ = B.bicgstabl(db, 2, dx, 1e-8, 1000, precision=dv.BKH_FP32_IR)
    print(f"cfg4 fp32ir csr-stream SpMV {info0.time_ms:.1f} ms")
    x_graph = dx.numpy()
    dx.upload(np.zeros(B.N))
    infop, retp, _ = B.bicgstabl(db, 2, dx, 1e-8, 1000, precision=dv.BKH_FP32_IR, path=2)
    assert dv.Batch.last_path() == 2
    print(f"cfg4 fp32ir persistent launch, csr-stream SpMV: {infop.time_ms:.1f} ms")
    assert np.array_equal(dx.numpy(), x_graph) and np.array_equal(retp, ret)
    B.spmv_kernel = kernel
    dx.upload(np.zeros(B.N))
    info64, ret64, tr64 = B.bicgstabl(db, 2, dx, 1e-8, 1000)
    print(f"cfg4 fp64 batched {info64.time_ms:.1f} ms max outer={info64.iters} max true res={tr64.max():.2e}")
    assert (ret64 >= 0).all() and tr64.max() < 1.5e-8
    x64 = dx.numpy()
    dx.upload(np.zeros(B.N))
    info64p, ret64p, _ = B.bicgstabl(db, 2, dx, 1e-8, 1000, path=2)
    print(f"cfg4 fp64 persistent launch {info64p.time_ms:.1f} ms")
    assert np.array_equal(dx.numpy(), x64) and np.array_equal(ret64p, ret64)
    # full-size parity on sampled systems (smallest, quantiles, largest): the oracle's
    # BiCGSTAB(2) (src/bicgstabl.c) to the same tol breaks within +-5 % of the batched
    # FP64 index, and x at the device's index matches the oracle's iterate there
    sizes = np.diff(B.offsets)
    order = np.argsort(sizes, kind="stable")
    picks = sorted({int(order[k]) for k in (0, len(order) // 4, len(order) // 2, 3 * len(order) // 4,
                                               len(order) - 1)})
    for q in picks:
        n, rp, ci, v = gen.random_rows(gen.cfg4_system_seed(q), int(sizes[q]))
        o = int(B.offsets[q])
        bq = b[o:o + n]
        cret, _ = po.bicgstabl(n, rp, ci, v, bq, 2, 1e-8, 1000)
        assert ret64[q] >= 0 and cret >= 0, (q, ret64[q], cret)
        count_index_mismatch(ret64[q], cret)
        xk = oracle_at_index(lambda t, k: po.bicgstabl(n, rp, ci, v, bq, 2, t, k), int(ret64[q]))
        if abs(int(ret64[q]) - cret) > max(1, int(0.05 * cret)):
            # a near-tie: the oracle's own trajectory meets tol at the device's index too,
            # and its residual at its break index sits just below tol (BiCGSTAB's residual
            # is not monotone, so a break missed by rounding can move by two iterations)
            def true_rel(xv):
                return np.linalg.norm(bq - po.csrgemv("n", n, rp, ci, v, xv)) / np.linalg.norm(bq)
            xc = oracle_at_index(lambda t, k: po.bicgstabl(n, rp, ci, v, bq, 2, t, k), cret)
            assert true_rel(xk) < 1e-8 and true_rel(xc) > 0.5e-8, (q, ret64[q], cret, true_rel(xk), true_rel(xc))
            print(f"cfg4 system {q}: near-tie at tol (oracle residual {true_rel(xc):.2e} at its break)")
        err = np.linalg.norm(x64[o:o + n] - xk) / np.linalg.norm(xk)
        print(f"cfg4 system {q}: n={n} device index {ret64[q]} oracle {cret} x rel diff {err:.2e}")
        assert err < 1e-8, (q, err)
