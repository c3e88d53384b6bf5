"""Device-level Python handles over the bksp_hip.h C-ABI.

``Matrix`` keeps a CSR resident in HBM, ``DeviceArray`` a device vector; the
solver wrappers return a ``BkhInfo`` with the bksp.h return value, iteration
index, recursive and FP64 true residuals and the device time of the solve.
"""
import ctypes

import numpy as np

from ._lib import BkhInfo, BkhOpts, load

BKH_FP64 = 0
BKH_FP32_IR = 1


class BkspDeviceError(RuntimeError):
    pass


def last_error():
    msg = load().bkh_last_error()
    return msg.decode() if msg else ""


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class DeviceArray:
    """A device buffer holding a NumPy-typed array (float64 or float32)."""

    def __init__(self, shape, dtype=np.float64):
        self.shape = tuple(np.atleast_1d(shape)) if not isinstance(shape, tuple) else shape
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.ptr = load().bkh_malloc(max(self.nbytes, 1))
        if not self.ptr:
            raise BkspDeviceError("bkh_malloc failed: " + last_error())

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        if load().bkh_h2d(self.ptr, _ptr(a), self.nbytes):
            raise BkspDeviceError("bkh_h2d failed: " + last_error())

    def zero(self):
        """Sets every byte to 0 on the device (a fill kernel; no host transfer)."""
        if load().bkh_memset(self.ptr, 0, self.nbytes):
            raise BkspDeviceError("bkh_memset failed: " + last_error())

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        if load().bkh_d2h(_ptr(out), self.ptr, self.nbytes):
            raise BkspDeviceError("bkh_d2h failed: " + last_error())
        return out

    def free(self):
        if self.ptr:
            load().bkh_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Matrix:
    """A device-resident CSR (bkh_matrix). ``symmetric=True`` means the arrays hold
    one stored triangle, as a bksp "symmetric" dspmat does."""

    def __init__(self, n, row_ptr, col, val, symmetric=False, transpose=False, fp32=False):
        self._rp = np.ascontiguousarray(row_ptr, dtype=np.int32)
        self._ci = np.ascontiguousarray(col, dtype=np.int32)
        self._v = np.ascontiguousarray(val, dtype=np.float64)
        self.n = int(n)
        self.handle = load().bkh_matrix_create(self.n, _ptr(self._rp), _ptr(self._ci), _ptr(self._v),
                                               int(symmetric), int(transpose), int(fp32))
        if not self.handle:
            raise BkspDeviceError("bkh_matrix_create failed: " + last_error())
        self.nnz = load().bkh_matrix_nnz(self.handle)
        self.dia_entries = load().bkh_matrix_dia_entries(self.handle)
        self.nblocks = load().bkh_matrix_blocks(self.handle)
        del self._rp, self._ci, self._v

    def destroy(self):
        if self.handle:
            load().bkh_matrix_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    def set_spmv(self, variant):
        """0: workgroup CSR-stream, 1: wave CSR-stream, 2: SELL-64 (16-bit column
        deltas when they fit), 3: SELL-64 with 32-bit columns."""
        if load().bkh_matrix_set_spmv(self.handle, int(variant)):
            raise BkspDeviceError("bkh_matrix_set_spmv failed")
        self.nblocks = load().bkh_matrix_blocks(self.handle)

    def tune(self, max_s=16):
        """bkh_matrix_tune: keeps the fastest SpMV variant / wave blocking and bkh_spmm's
        columns per pass for s = 1..max_s; returns what it chose and measured."""
        from ._lib import BkhMatrixTuning
        r = BkhMatrixTuning()
        rc = load().bkh_matrix_tune(self.handle, int(max_s), ctypes.byref(r))
        if rc:
            raise BkspDeviceError(f"bkh_matrix_tune -> {rc}: {last_error()}")
        self.nblocks = load().bkh_matrix_blocks(self.handle)
        return {"spmv_variant": r.spmv_variant, "rows_per_wave": r.rows_per_wave,
                "nnz_per_wave": r.nnz_per_wave, "candidates": r.candidates,
                "spmv_ms": r.spmv_ms, "spmv_ms_default": r.spmv_ms_default,
                "spmm_width": {s: r.spmm_width[s] for s in range(1, max_s + 1)},
                "spmm_ms": {s: r.spmm_ms[s] for s in range(1, max_s + 1)},
                "spmm_ms_full": {s: r.spmm_ms_full[s] for s in range(1, max_s + 1)}}

    def spmv(self, d_x, d_y, op="n", alpha=1.0, beta=0.0):
        rc = load().bkh_spmv(self.handle, op.encode(), alpha, d_x.ptr, beta, d_y.ptr)
        if rc:
            raise BkspDeviceError(f"bkh_spmv -> {rc}: {last_error()}")

    def spmm(self, s, d_X, d_Y, alpha=1.0, beta=0.0):
        """Y = alpha A X + beta Y for s column-major columns of n entries."""
        rc = load().bkh_spmm(self.handle, s, alpha, d_X.ptr, beta, d_Y.ptr)
        if rc:
            raise BkspDeviceError(f"bkh_spmm -> {rc}: {last_error()}")

    def spmm_gram(self, s, d_P, d_AP):
        """AP = A P and the Gram P^T A P (s x s numpy array) for s column-major columns,
        by block CG's path; returns (G, path) with path 0 (separate SpMM and Gram),
        1 (fused kernel) or 2 (pipelined slice-diagonal fused kernel)."""
        G = np.zeros(s * s)
        path = ctypes.c_int(-1)
        rc = load().bkh_spmm_gram(self.handle, s, d_P.ptr, d_AP.ptr, G.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                  ctypes.byref(path))
        if rc:
            raise BkspDeviceError(f"bkh_spmm_gram -> {rc}: {last_error()}")
        return G.reshape(s, s, order="F"), path.value

    def spmv_bytes(self, fp32=False):
        """Bytes one SpMV with the current kernel moves (format bytes)."""
        return load().bkh_matrix_spmv_bytes(self.handle, int(fp32))

    def bench_cg_kernel(self, reps=20, fp32=False):
        """Average ms of CG's fused SpMV + direction-update kernel."""
        ms = ctypes.c_double()
        _check(load().bkh_bench_cg_kernel(self.handle, int(fp32), reps, ctypes.byref(ms)), "bkh_bench_cg_kernel")
        return ms.value

    def spmv32(self, d_x, d_y):
        rc = load().bkh_spmv32(self.handle, d_x.ptr, d_y.ptr)
        if rc:
            raise BkspDeviceError(f"bkh_spmv32 -> {rc}: {last_error()}")


def make_opts(precision=BKH_FP64, graph_iters=0, fix_blcg_p0=0, max_inner=0, inner_tol=0.0, path=0):
    return BkhOpts(precision, graph_iters, fix_blcg_p0, max_inner, inner_tol, path)


def _check(rc, name):
    if rc == -5 or rc == -6:
        raise BkspDeviceError(f"{name} -> {rc}: {last_error()}")


def cg(M, d_b, d_x, tol, max_iter, **kw):
    opts = make_opts(**kw)
    info = BkhInfo()
    rc = load().bkh_cg(M.handle, d_b.ptr, d_x.ptr, tol, max_iter, ctypes.byref(opts), ctypes.byref(info))
    _check(rc, "bkh_cg")
    return info


def _solve1(fn, name, M, d_b, d_x, tol, max_iter, **kw):
    opts = make_opts(**kw)
    info = BkhInfo()
    rc = fn(M.handle, d_b.ptr, d_x.ptr, tol, max_iter, ctypes.byref(opts), ctypes.byref(info))
    _check(rc, name)
    return info


def cr(M, d_b, d_x, tol, max_iter, **kw):
    return _solve1(load().bkh_cr, "bkh_cr", M, d_b, d_x, tol, max_iter, **kw)


def bicg(M, d_b, d_x, tol, max_iter, **kw):
    return _solve1(load().bkh_bicg, "bkh_bicg", M, d_b, d_x, tol, max_iter, **kw)


def bicgstabl(M, d_b, l, d_x, tol, max_iter, **kw):
    opts = make_opts(**kw)
    info = BkhInfo()
    rc = load().bkh_bicgstabl(M.handle, d_b.ptr, l, d_x.ptr, tol, max_iter, ctypes.byref(opts),
                              ctypes.byref(info))
    _check(rc, "bkh_bicgstabl")
    return info


class ZMatrix:
    """A device-resident complex general CSR (bkh_zmatrix); values complex128."""

    def __init__(self, n, row_ptr, col, val):
        rp = np.ascontiguousarray(row_ptr, dtype=np.int32)
        ci = np.ascontiguousarray(col, dtype=np.int32)
        v = np.ascontiguousarray(val, dtype=np.complex128)
        self.n = int(n)
        self.handle = load().bkh_zmatrix_create(self.n, _ptr(rp), _ptr(ci), _ptr(v))
        if not self.handle:
            raise BkspDeviceError("bkh_zmatrix_create failed: " + last_error())
        self.dia_entries = load().bkh_zmatrix_dia_entries(self.handle)
        self.split_rows = load().bkh_zmatrix_split_rows(self.handle)

    def destroy(self):
        if self.handle:
            load().bkh_zmatrix_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    def spmv(self, d_x, d_y, op="n"):
        """y = A x (op 'n') or A^H x (op 't')."""
        rc = load().bkh_zspmv(self.handle, op.encode(), d_x.ptr, d_y.ptr)
        if rc:
            raise BkspDeviceError(f"bkh_zspmv -> {rc}: {last_error()}")

    def bench_spmv(self, reps=20, op="n"):
        """Average ms of one complex SpMV launch (HIP events, back-to-back launches)."""
        ms = ctypes.c_double()
        _check(load().bkh_zspmv_bench(self.handle, op.encode(), reps, ctypes.byref(ms)), "bkh_zspmv_bench")
        return ms.value

    def spmm1(self, s, d_X, d_Y, alpha=1.0, beta=0.0, op="n"):
        """Y = beta Y + alpha op(A) X for s column-major columns."""
        al = np.array([alpha], np.complex128)
        be = np.array([beta], np.complex128)
        rc = load().bkh_zspmm1(self.handle, op.encode(), _ptr(al), s, d_X.ptr, _ptr(be), d_Y.ptr)
        if rc:
            raise BkspDeviceError(f"bkh_zspmm1 -> {rc}: {last_error()}")


def zbicg(M, d_b, d_x, tol, max_iter):
    info = BkhInfo()
    rc = load().bkh_zbicg(M.handle, d_b.ptr, d_x.ptr, tol, max_iter, ctypes.byref(info))
    _check(rc, "bkh_zbicg")
    return info


def zbicgstabl(M, d_b, l, d_x, tol, max_iter):
    info = BkhInfo()
    rc = load().bkh_zbicgstabl(M.handle, d_b.ptr, l, d_x.ptr, tol, max_iter, ctypes.byref(info))
    _check(rc, "bkh_zbicgstabl")
    return info


def zblbicgstabl(M, s, d_B, l, d_X, tol, max_iter):
    """Complex block BiCGSTAB(l); d_B, d_X hold m x s column-major complex blocks."""
    info = BkhInfo()
    rc = load().bkh_zblbicgstabl(M.handle, s, d_B.ptr, l, d_X.ptr, tol, max_iter, ctypes.byref(info))
    _check(rc, "bkh_zblbicgstabl")
    return info


def zqr(d_V, m, s):
    """Thin QR of an m x s complex panel (column-major) in place; returns R."""
    R = np.zeros((s, s), np.complex128, order="F")
    rc = load().bkh_zqr(m, s, d_V.ptr, R.ctypes.data_as(ctypes.c_void_p))
    if rc:
        raise BkspDeviceError(f"bkh_zqr -> {rc}: {last_error()}")
    return R


def zshbicgstabl(M, d_b, l, sigma, d_x, tol, max_iter):
    """Complex shifted BiCGSTAB(l); d_x holds n*(len(sigma)+1) complex entries (seed
    first). Returns (info, per-system FP64 true residuals)."""
    info = BkhInfo()
    sig = np.ascontiguousarray(sigma, dtype=np.complex128)
    tr = np.zeros(len(sig) + 1)
    rc = load().bkh_zshbicgstabl(M.handle, d_b.ptr, l, sig.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                 len(sig), d_x.ptr, tol, max_iter, ctypes.byref(info),
                                 tr.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    _check(rc, "bkh_zshbicgstabl")
    return info, tr


def blcg(M, s, d_B, d_X, tol, max_iter, **kw):
    """Block CG; d_B, d_X hold m x s column-major blocks (Fortran order)."""
    opts = make_opts(**kw)
    info = BkhInfo()
    rc = load().bkh_blcg(M.handle, s, d_B.ptr, d_X.ptr, tol, max_iter, ctypes.byref(opts),
                         ctypes.byref(info))
    _check(rc, "bkh_blcg")
    return info


def blbicgstabl(M, s, d_B, l, d_X, tol, max_iter, **kw):
    """Block BiCGSTAB(l); d_B, d_X hold m x s column-major blocks (Fortran order)."""
    opts = make_opts(**kw)
    info = BkhInfo()
    rc = load().bkh_blbicgstabl(M.handle, s, d_B.ptr, l, d_X.ptr, tol, max_iter, ctypes.byref(opts),
                                ctypes.byref(info))
    _check(rc, "bkh_blbicgstabl")
    return info


def shbicgstabl(M, d_b, l, sigma, d_x, tol, max_iter, **kw):
    """Shifted BiCGSTAB(l); d_x holds n*(len(sigma)+1) entries (seed first).
    Returns (info, per-system FP64 true residuals)."""
    opts = make_opts(**kw)
    info = BkhInfo()
    sig = np.ascontiguousarray(sigma, dtype=np.float64)
    tr = np.zeros(len(sig) + 1)
    rc = load().bkh_shbicgstabl(M.handle, d_b.ptr, l, sig.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                len(sig), d_x.ptr, tol, max_iter, ctypes.byref(opts), ctypes.byref(info),
                                tr.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    _check(rc, "bkh_shbicgstabl")
    return info, tr


def dot(d_x, d_y):
    out = ctypes.c_double()
    rc = load().bkh_dot(int(np.prod(d_x.shape)), d_x.ptr, d_y.ptr, ctypes.byref(out))
    _check(rc, "bkh_dot")
    return out.value


def device_name():
    buf = ctypes.create_string_buffer(256)
    if load().bkh_device_name(buf, 256):
        raise BkspDeviceError(last_error())
    return buf.value.decode()


def bench_spmv(M, d_x, d_y, reps=50, fp32=False):
    """Average ms per SpMV launch (HIP events on the kernels' stream)."""
    out = ctypes.c_double()
    rc = load().bkh_bench_spmv(M.handle, int(fp32), d_x.ptr, d_y.ptr, reps, ctypes.byref(out))
    _check(rc, "bkh_bench_spmv")
    return out.value


def bench_copy(d_a, d_b, nbytes, reps=20):
    out = ctypes.c_double()
    rc = load().bkh_bench_copy(d_a.ptr, d_b.ptr, nbytes, reps, ctypes.byref(out))
    _check(rc, "bkh_bench_copy")
    return out.value


class Batch:
    """Independent CSR systems stacked block-diagonally on the device (bkh_batch).
    ``systems`` is a list of (n, row_ptr, col, val). Vectors of the batch are
    concatenated: system s occupies rows offset(s) .. offset(s + 1)."""

    def __init__(self, systems, fp32=False):
        ns = len(systems)
        self.ns = ns
        self._keep = [(np.ascontiguousarray(rp, np.int32), np.ascontiguousarray(ci, np.int32),
                       np.ascontiguousarray(v, np.float64)) for (_, rp, ci, v) in systems]
        n_each = np.array([int(sy[0]) for sy in systems], np.int32)
        IP = ctypes.POINTER(ctypes.c_int)
        DP = ctypes.POINTER(ctypes.c_double)
        rps = (IP * ns)(*[k[0].ctypes.data_as(IP) for k in self._keep])
        cis = (IP * ns)(*[k[1].ctypes.data_as(IP) for k in self._keep])
        vs = (DP * ns)(*[k[2].ctypes.data_as(DP) for k in self._keep])
        self.handle = load().bkh_batch_create(ns, n_each.ctypes.data_as(IP), rps, cis, vs, int(fp32))
        del self._keep
        if not self.handle:
            raise BkspDeviceError("bkh_batch_create failed: " + last_error())
        self.N = load().bkh_batch_rows(self.handle)
        self.nnz = load().bkh_batch_nnz(self.handle)
        self.offsets = np.array([load().bkh_batch_offset(self.handle, q) for q in range(ns + 1)], np.int64)

    def destroy(self):
        if self.handle:
            load().bkh_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    def layout(self):
        a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        load().bkh_batch_layout(self.handle, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
        return a.value, b.value, c.value

    @property
    def spmv_kernel(self):
        """1: the column-chunked SpMV, 0: the CSR-stream units."""
        return load().bkh_batch_spmv_kernel(self.handle)

    @spmv_kernel.setter
    def spmv_kernel(self, chunked):
        if load().bkh_batch_set_spmv_kernel(self.handle, int(chunked)):
            raise ValueError("this batch has no column-chunked copy (some row's columns do not ascend)")

    def tune(self, fp32=False):
        ms = ctypes.c_double()
        rc = load().bkh_batch_tune(self.handle, int(fp32), ctypes.byref(ms))
        _check(rc, "bkh_batch_tune")
        return ms.value

    def spmv(self, d_x, d_y, fp32=False):
        _check(load().bkh_batch_spmv(self.handle, int(fp32), d_x.ptr, d_y.ptr), "bkh_batch_spmv")

    def bicgstabl(self, d_b, l, d_x, tol, max_iter, **kw):
        """Returns (info, per-system return values, per-system FP64 true residuals)."""
        opts = make_opts(**kw)
        info = BkhInfo()
        ret = np.zeros(self.ns, np.int32)
        tr = np.zeros(self.ns)
        rc = load().bkh_batch_bicgstabl(self.handle, d_b.ptr, l, d_x.ptr, tol, max_iter, ctypes.byref(opts),
                                        ret.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                        tr.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(info))
        if rc < -4:
            raise BkspDeviceError(f"bkh_batch_bicgstabl -> {rc}: {last_error()}")
        return info, ret, tr

    @staticmethod
    def last_path():
        """1: the last batched solve ran per-phase graphs; 2: one persistent launch."""
        return load().bkh_batch_last_path()


def pcg(M, kind, d_b, d_x, tol, max_iter, **kw):
    """Preconditioned CG: kind 0 = none, 1 = Jacobi."""
    opts = make_opts(**kw)
    info = BkhInfo()
    _check(load().bkh_pcg(M.handle, int(kind), d_b.ptr, d_x.ptr, tol, max_iter, ctypes.byref(opts),
                          ctypes.byref(info)), "bkh_pcg")
    return info


class FsaiModel:
    """The learned-FSAI edge MLP (FP32 master weights on the device, Adam state)."""

    def __init__(self, params=None):
        self.nparam = load().bkh_fsai_nparam()
        p = None
        if params is not None:
            self._p = np.ascontiguousarray(params, np.float32)
            assert self._p.size == self.nparam
            p = self._p.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        self.handle = load().bkh_fsai_model_create(p)
        if not self.handle:
            raise BkspDeviceError("bkh_fsai_model_create failed: " + last_error())

    def get(self):
        out = np.empty(self.nparam, np.float32)
        _check(load().bkh_fsai_model_get(self.handle, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))), "get")
        return out

    def set(self, params):
        p = np.ascontiguousarray(params, np.float32)
        _check(load().bkh_fsai_model_set(self.handle, p.ctypes.data_as(ctypes.POINTER(ctypes.c_float))), "set")

    def destroy(self):
        if self.handle:
            load().bkh_fsai_model_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class FsaiField:
    """tril(A) of an SPD matrix (with probes Z, P x n, for training) on the device."""

    def __init__(self, n, rpl, cil, vl, Z=None):
        rpl = np.ascontiguousarray(rpl, np.int32)
        cil = np.ascontiguousarray(cil, np.int32)
        vl = np.ascontiguousarray(vl, np.float64)
        P = 0 if Z is None else int(Z.shape[0])
        Zc = np.ascontiguousarray(Z if Z is not None else np.zeros(1), np.float64)
        self.n, self.nnzl = int(n), int(rpl[-1])
        self.handle = load().bkh_fsai_field_create(self.n, _ptr(rpl), _ptr(cil), _ptr(vl), P, _ptr(Zc))
        if not self.handle:
            raise BkspDeviceError("bkh_fsai_field_create failed: " + last_error())

    def factor(self, model, want=False):
        g = np.empty(self.nnzl) if want else None
        _check(load().bkh_fsai_factor(model.handle, self.handle, _ptr(g) if want else None), "bkh_fsai_factor")
        return g

    def loss_grad(self, model):
        L = ctypes.c_double()
        g = np.empty(load().bkh_fsai_nparam())
        _check(load().bkh_fsai_loss_grad(model.handle, self.handle, ctypes.byref(L), _ptr(g)), "loss_grad")
        return L.value, g

    def train_step(self, model, lr, want_loss=False):
        L = ctypes.c_double()
        _check(load().bkh_fsai_train_step(model.handle, self.handle, lr, ctypes.byref(L) if want_loss else None),
               "train_step")
        return L.value if want_loss else None

    def bench_forward(self, model, reps=10):
        ms = ctypes.c_double()
        _check(load().bkh_fsai_bench_forward(model.handle, self.handle, reps, ctypes.byref(ms)), "bench_forward")
        return ms.value

    def pcg(self, d_b, d_x, tol, max_iter, **kw):
        opts = make_opts(**kw)
        info = BkhInfo()
        _check(load().bkh_fsai_pcg(self.handle, d_b.ptr, d_x.ptr, tol, max_iter, ctypes.byref(opts),
                                   ctypes.byref(info)), "bkh_fsai_pcg")
        return info

    def pcg_plain(self, kind, d_b, d_x, tol, max_iter, **kw):
        """PCG on this field's matrix with kind 0 (none) or 1 (Jacobi)."""
        opts = make_opts(**kw)
        info = BkhInfo()
        A = load().bkh_fsai_field_matrix(self.handle)
        _check(load().bkh_pcg(A, int(kind), d_b.ptr, d_x.ptr, tol, max_iter, ctypes.byref(opts),
                              ctypes.byref(info)), "bkh_pcg")
        return info

    def destroy(self):
        if self.handle:
            load().bkh_fsai_field_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass
