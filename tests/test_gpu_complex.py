"""GPU parity of the complex FP64 family (bkh_z*, and the drop-in zbicg /
zbicgstabl / zshbicgstabl / zcsrgemm0 / zcsrgemm1 symbols) against the oracle's
restatement of src/bicg.c:91-177, src/bicgstabl.c:141-276, src/shbicgstabl.c:288-592
and src/matvec.c:55-98, 174-217.

The CSR products sum every row in the reference's order with its C99 complex
arithmetic (no FMA), so they are BITWISE equal to the oracle. The solvers' dot
products are tree-summed on the device (the oracle sums in index order), so
trajectories agree to rounding: the break index within 5 % (at least 1), the
solution to 1e-8 relative of the oracle's iterate at the device's break index, and the reported FP64 true
residual equal to the host's."""
import ctypes

import numpy as np
import pytest

from bksp_amd import device as dv
from bksp_amd import load
from conftest import count_index_mismatch, oracle_at_index, zcsr_dense, zsystem
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def close_index(a, b):
    return abs(a - b) <= max(1, int(0.05 * max(a, b)))


def zrand(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


@pytest.mark.parametrize("m", [8, 48])
def test_zspmv_bitwise(m):
    n, rp, ci, v = zsystem(m, seed=m)
    M = dv.ZMatrix(n, rp, ci, v)
    x = zrand(np.random.default_rng(1), n)
    dx, dy = dv.DeviceArray.from_numpy(x), dv.DeviceArray((n,), np.complex128)
    for op in "nt":
        M.spmv(dx, dy, op)
        assert np.array_equal(dy.numpy(), po.zcsrgemm0(op, n, rp, ci, v, x)[:, 0]), op


def test_zspmm1_bitwise():
    n, rp, ci, v = zsystem(24, seed=2)
    M = dv.ZMatrix(n, rp, ci, v)
    rng = np.random.default_rng(3)
    X, C = zrand(rng, n, 3), zrand(rng, n, 3)
    dX = dv.DeviceArray.from_numpy(X.ravel(order="F"))
    for op in "nt":
        dC = dv.DeviceArray.from_numpy(C.ravel(order="F"))
        M.spmm1(3, dX, dC, 0.7 - 0.4j, -1.25 + 0.5j, op)
        want = po.zcsrgemm1(op, 0.7 - 0.4j, n, rp, ci, v, X, -1.25 + 0.5j, C)
        assert np.array_equal(dC.numpy().reshape(3, n).T, want), op


def gpu_solve(kind, n, rp, ci, v, b, *args):
    M = dv.ZMatrix(n, rp, ci, v)
    db = dv.DeviceArray.from_numpy(b)
    if kind == "shifted":
        l, sig, tol, mi = args
        dx = dv.DeviceArray((n * (len(sig) + 1),), np.complex128)
        dx.zero()
        info, tr = dv.zshbicgstabl(M, db, l, sig, dx, tol, mi)
        return info, dx.numpy().reshape(len(sig) + 1, n), tr
    dx = dv.DeviceArray((n,), np.complex128)
    dx.zero()
    if kind == "bicg":
        info = dv.zbicg(M, db, dx, *args)
    else:
        l, tol, mi = args
        info = dv.zbicgstabl(M, db, l, dx, tol, mi)
    return info, dx.numpy(), None


def test_zbicg_parity():
    n, rp, ci, v = zsystem(40, seed=4)
    A = zcsr_dense(n, rp, ci, v)
    b = zrand(np.random.default_rng(5), n)
    info, x, _ = gpu_solve("bicg", n, rp, ci, v, b, 1e-10, 3000)
    cret, cx = po.zbicg(n, rp, ci, v, b, 1e-10, 3000)
    assert info.ret >= 0 and cret >= 0 and close_index(info.ret, cret), (info.ret, cret)
    res = np.linalg.norm(b - A @ x) / np.linalg.norm(b)
    assert res < 1e-9 and abs(info.true_res - res) <= 1e-6 * res, (res, info.true_res)
    count_index_mismatch(info.ret, cret)
    xk = oracle_at_index(lambda t, k: po.zbicg(n, rp, ci, v, b, t, k), info.ret)
    assert rel(x, xk) < 1e-8


@pytest.mark.parametrize("l", [1, 2, 4])
def test_zbicgstabl_parity(l):
    n, rp, ci, v = zsystem(40, seed=6)
    A = zcsr_dense(n, rp, ci, v)
    b = zrand(np.random.default_rng(7), n)
    info, x, _ = gpu_solve("bicgstabl", n, rp, ci, v, b, l, 1e-10, 2000)
    cret, cx = po.zbicgstabl(n, rp, ci, v, b, l, 1e-10, 2000)
    assert info.ret >= 0 and cret >= 0 and close_index(info.ret, cret), (l, info.ret, cret)
    res = np.linalg.norm(b - A @ x) / np.linalg.norm(b)
    assert res < 1e-9 and abs(info.true_res - res) <= 1e-6 * res
    count_index_mismatch(info.ret, cret)
    xk = oracle_at_index(lambda t, k: po.zbicgstabl(n, rp, ci, v, b, l, t, k), info.ret)
    assert rel(x, xk) < 1e-8


@pytest.mark.parametrize("l", [1, 2, 4])
def test_zshbicgstabl_parity(l):
    n, rp, ci, v = zsystem(32, seed=8)
    A = zcsr_dense(n, rp, ci, v)
    b = zrand(np.random.default_rng(9), n)
    sig = np.array([0.02 + 0.1j, 0.1, 0.5 - 0.2j, 1.0 + 1.0j, 2.0, 0.3 + 0.7j])
    info, X, tr = gpu_solve("shifted", n, rp, ci, v, b, l, sig, 1e-10, 1000)
    cret, CX = po.zshbicgstabl(n, rp, ci, v, b, l, sig, 1e-10, 1000)
    assert info.ret >= 0 and cret >= 0 and close_index(info.ret, cret), (l, info.ret, cret)
    for k in range(len(sig) + 1):
        s = 0.0 if k == 0 else sig[k - 1]
        res = np.linalg.norm(b - A @ X[k] - s * X[k]) / np.linalg.norm(b)
        assert abs(tr[k] - res) <= 1e-6 * res + 1e-300, (k, tr[k], res)
        assert res < (1e-10 if k == 0 else 1e-8), (l, k, res)
    count_index_mismatch(info.ret, cret)
    Xk = oracle_at_index(lambda t, k: po.zshbicgstabl(n, rp, ci, v, b, l, sig, t, k), info.ret)
    assert rel(X, Xk) < 1e-8


def arrow_zsystem(m, seed):
    """zsystem(m) plus a dense first row and first column (seeded complex entries of
    scale 10, the diagonal being about 7e3): one row of A and one of A^T far longer than
    their slice's other rows."""
    import scipy.sparse as sp
    n, rp, ci, v = zsystem(m, seed=seed)
    rng = np.random.default_rng(seed + 100)
    D = sp.lil_matrix((n, n), dtype=np.complex128)
    D[0, 1:] = 10.0 * zrand(rng, n - 1)
    D[1:, 0] = (10.0 * zrand(rng, n - 1))[:, None]
    B = (zcsr_dense(n, rp, ci, v) + D.tocsr()).tocsr()
    B.sort_indices()
    return n, B.indptr.astype(np.int32), B.indices.astype(np.int32), B.data.astype(np.complex128)


def test_dense_row_and_column_split_bitwise():
    """A row longer than its slice's SELL width keeps its head in the slice and the rest
    in a CSR tail summed after it in stored order: still bitwise the oracle's products,
    and the solvers (no slice-diagonal copy for such a matrix) still match the oracle."""
    n, rp, ci, v = arrow_zsystem(40, seed=6)
    M = dv.ZMatrix(n, rp, ci, v)
    assert M.split_rows == 2 and M.dia_entries == 0, (M.split_rows, M.dia_entries)
    rng = np.random.default_rng(22)
    x = zrand(rng, n)
    dx, dy = dv.DeviceArray.from_numpy(x), dv.DeviceArray((n,), np.complex128)
    for op in "nt":
        M.spmv(dx, dy, op)
        assert np.array_equal(dy.numpy(), po.zcsrgemm0(op, n, rp, ci, v, x)[:, 0]), op
    X, C = zrand(rng, n, 3), zrand(rng, n, 3)
    dX = dv.DeviceArray.from_numpy(X.ravel(order="F"))
    for op in "nt":
        dC = dv.DeviceArray.from_numpy(C.ravel(order="F"))
        M.spmm1(3, dX, dC, 0.7 - 0.4j, -1.25 + 0.5j, op)
        want = po.zcsrgemm1(op, 0.7 - 0.4j, n, rp, ci, v, X, -1.25 + 0.5j, C)
        assert np.array_equal(dC.numpy().reshape(3, n).T, want), op
    A = zcsr_dense(n, rp, ci, v)
    b = zrand(np.random.default_rng(23), n)
    for kind, args, ora in (("bicg", (1e-10, 3000), lambda t, k: po.zbicg(n, rp, ci, v, b, t, k)),
                            ("bicgstabl", (2, 1e-10, 2000),
                             lambda t, k: po.zbicgstabl(n, rp, ci, v, b, 2, t, k))):
        info, x, _ = gpu_solve(kind, n, rp, ci, v, b, *args)
        cret, cx = ora(1e-10, args[-1])
        assert info.ret >= 0 and cret >= 0 and close_index(info.ret, cret), (kind, info.ret, cret)
        res = np.linalg.norm(b - A @ x) / np.linalg.norm(b)
        assert res < 1e-9 and abs(info.true_res - res) <= 1e-6 * res, (kind, res, info.true_res)
        assert rel(x, oracle_at_index(ora, info.ret)) < 1e-8, kind


def test_real_data_matches_real_solver():
    """With zero imaginary parts the complex BiCGSTAB(l) performs the real one's
    arithmetic: same break index as the real device solver and the oracle."""
    from bksp_amd import gen
    n, rp, ci, v = gen.convdiff2d(32, 10.0, 20.0)
    b = np.random.default_rng(10).standard_normal(n)
    info, x, _ = gpu_solve("bicgstabl", n, rp, ci, v.astype(complex), b.astype(complex), 2, 1e-10, 500)
    ret, rx = po.bicgstabl(n, rp, ci, v, b, 2, 1e-10, 500)
    assert info.ret == ret and np.abs(x.imag).max() == 0.0 and rel(x.real, rx) < 1e-9


def test_max_iter_zero_and_bad_l():
    n, rp, ci, v = zsystem(8, seed=11)
    b = zrand(np.random.default_rng(12), n)
    info, _, _ = gpu_solve("bicgstabl", n, rp, ci, v, b, 2, 1e-10, 0)
    assert info.ret == 0
    M = dv.ZMatrix(n, rp, ci, v)
    db, dx = dv.DeviceArray.from_numpy(b), dv.DeviceArray((n,), np.complex128)
    assert load().bkh_zbicgstabl(M.handle, db.ptr, 0, dx.ptr, 1e-10, 10, None) == -6   # BKH_BAD_ARGUMENT


class ZSpMat(ctypes.Structure):                 # include/bksp.h zspmat
    _fields_ = [("row", ctypes.POINTER(ctypes.c_int)), ("col", ctypes.POINTER(ctypes.c_int)),
                ("nnz", ctypes.POINTER(ctypes.c_int)), ("row_size", ctypes.POINTER(ctypes.c_int)),
                ("col_size", ctypes.POINTER(ctypes.c_int)), ("value", ctypes.c_void_p),
                ("format", ctypes.c_char_p), ("type", ctypes.c_char_p)]


def test_dropin_zsolvers_and_kernels():
    n, rp, ci, v = zsystem(24, seed=13)
    A = zcsr_dense(n, rp, ci, v)
    rp = np.ascontiguousarray(rp, np.int32); ci = np.ascontiguousarray(ci, np.int32)
    v = np.ascontiguousarray(v, np.complex128)
    ints = [np.array([x], np.int32) for x in (len(v), n, n)]
    IP = ctypes.POINTER(ctypes.c_int)
    Z = ZSpMat(rp.ctypes.data_as(IP), ci.ctypes.data_as(IP), ints[0].ctypes.data_as(IP),
               ints[1].ctypes.data_as(IP), ints[2].ctypes.data_as(IP), v.ctypes.data, b"csr", b"unsymmetric")
    lib = load()
    b = zrand(np.random.default_rng(14), n)
    vp = lambda a: ctypes.c_void_p(a.ctypes.data)
    x = np.zeros(n, np.complex128)
    lib.zbicg.restype = ctypes.c_int
    lib.zbicg.argtypes = [ctypes.POINTER(ZSpMat), ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.c_void_p]
    ret = lib.zbicg(ctypes.byref(Z), vp(b), 1e-10, 2000, vp(x))
    cret, cx = po.zbicg(n, rp, ci, v, b, 1e-10, 2000)
    assert ret >= 0 and close_index(ret, cret) and np.linalg.norm(b - A @ x) < 1e-9 * np.linalg.norm(b)
    x = np.zeros(n, np.complex128)
    lib.zbicgstabl.restype = ctypes.c_int
    lib.zbicgstabl.argtypes = [ctypes.POINTER(ZSpMat), ctypes.c_void_p, ctypes.c_int, ctypes.c_double,
                               ctypes.c_int, ctypes.c_void_p]
    ret = lib.zbicgstabl(ctypes.byref(Z), vp(b), 2, 1e-10, 2000, vp(x))
    assert ret >= 0 and np.linalg.norm(b - A @ x) < 1e-9 * np.linalg.norm(b)
    sig = np.array([0.1 + 0.1j, 1.0], np.complex128)
    xs = np.zeros(3 * n, np.complex128)
    lib.zshbicgstabl.restype = ctypes.c_int
    lib.zshbicgstabl.argtypes = [ctypes.POINTER(ZSpMat), ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                 ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_void_p]
    ret = lib.zshbicgstabl(ctypes.byref(Z), vp(b), 2, vp(sig), 2, 1e-10, 1000, vp(xs))
    X = xs.reshape(3, n)
    assert ret >= 0
    for k, s in enumerate([0.0] + list(sig)):
        assert np.linalg.norm(b - A @ X[k] - s * X[k]) < 1e-8 * np.linalg.norm(b), k
    # the exported kernel symbols: zcsrgemm0 (0-based) and zcsrgemm1 (1-based)
    B = np.asfortranarray(zrand(np.random.default_rng(15), n, 2))
    C = np.zeros((n, 2), np.complex128, order="F")
    lib.zcsrgemm0.restype = ctypes.c_int
    assert lib.zcsrgemm0(ctypes.c_char(b"t"), n, 2, vp(v), vp(rp), vp(ci), len(v), vp(B), vp(C)) == 0
    assert np.array_equal(C, po.zcsrgemm0("t", n, rp, ci, v, B))
    C0 = np.asfortranarray(zrand(np.random.default_rng(16), n, 2))
    C1 = C0.copy(order="F")
    al, be = np.array([0.5 + 0.25j]), np.array([-1.0 + 0.5j])
    rp1, ci1 = rp + 1, ci + 1
    lib.zcsrgemm1.restype = ctypes.c_int
    assert lib.zcsrgemm1(ctypes.c_char(b"n"), vp(al), n, 2, vp(v), vp(rp1), vp(ci1), len(v), vp(B), vp(be),
                         vp(C1)) == 0
    assert np.array_equal(C1, po.zcsrgemm1("n", al[0], n, rp, ci, v, B, be[0], C0))


def test_zqr_invariants():
    """CholeskyQR2 on the device: Q^H Q = I, Q R = V, R upper triangular with a real
    positive diagonal, equal to LAPACK's R up to the sign of each row."""
    m, s = 500, 6
    V = zrand(np.random.default_rng(17), m, s)
    dV = dv.DeviceArray.from_numpy(V.ravel(order="F"))
    R = dv.zqr(dV, m, s)
    Q = dV.numpy().reshape(s, m).T
    assert np.allclose(Q.conj().T @ Q, np.eye(s), atol=1e-13)
    assert np.allclose(Q @ R, V, atol=1e-12)
    assert np.allclose(np.tril(R, -1), 0) and np.all(np.diag(R).real > 0) and np.all(np.diag(R).imag == 0)
    _, Rn = po.zqreco(V)
    d = np.sign(np.diag(Rn).real)
    assert np.allclose(R, d[:, None] * Rn, atol=1e-12)


@pytest.mark.parametrize("l,s", [(1, 3), (2, 4), (4, 2)])
def test_zblbicgstabl_parity(l, s):
    n, rp, ci, v = zsystem(32, seed=18)
    A = zcsr_dense(n, rp, ci, v)
    B = zrand(np.random.default_rng(19), n, s)
    M = dv.ZMatrix(n, rp, ci, v)
    dB = dv.DeviceArray.from_numpy(np.asfortranarray(B).ravel(order="F"))
    dX = dv.DeviceArray((n * s,), np.complex128)
    dX.zero()
    info = dv.zblbicgstabl(M, s, dB, l, dX, 1e-10, 500)
    X = dX.numpy().reshape(s, n).T
    cret, CX, ctr = po.zblbicgstabl(n, rp, ci, v, B, l, 1e-10, 500)
    assert info.ret >= 0 and cret >= 0 and close_index(info.ret, cret), (l, s, info.ret, cret)
    res = np.linalg.norm(B - A @ X) / np.linalg.norm(B)
    assert res < 1e-8 and abs(info.true_res - res) <= 1e-6 * res, (res, info.true_res)
    count_index_mismatch(info.ret, cret)
    Xk = oracle_at_index(lambda t, k: po.zblbicgstabl(n, rp, ci, v, B, l, t, k), info.ret)
    assert rel(X, Xk) < 1e-8


class ZMat(ctypes.Structure):                    # include/bksp.h zmat
    _fields_ = [("row_size", ctypes.c_int), ("col_size", ctypes.c_int), ("value", ctypes.c_void_p)]


def test_dropin_zblbicgstabl_and_zqreco(capfd):
    n, rp, ci, v = zsystem(20, seed=20)
    A = zcsr_dense(n, rp, ci, v)
    rp = np.ascontiguousarray(rp, np.int32); ci = np.ascontiguousarray(ci, np.int32)
    v = np.ascontiguousarray(v, np.complex128)
    ints = [np.array([x], np.int32) for x in (len(v), n, n)]
    IP = ctypes.POINTER(ctypes.c_int)
    Z = ZSpMat(rp.ctypes.data_as(IP), ci.ctypes.data_as(IP), ints[0].ctypes.data_as(IP),
               ints[1].ctypes.data_as(IP), ints[2].ctypes.data_as(IP), v.ctypes.data, b"csr", b"unsymmetric")
    B = np.asfortranarray(zrand(np.random.default_rng(21), n, 3))
    X = np.zeros((n, 3), np.complex128, order="F")
    Bm, Xm = ZMat(n, 3, B.ctypes.data), ZMat(n, 3, X.ctypes.data)
    lib = load()
    lib.zblbicgstabl.restype = ctypes.c_int
    lib.zblbicgstabl.argtypes = [ctypes.POINTER(ZSpMat), ctypes.POINTER(ZMat), ctypes.c_int, ctypes.c_double,
                                 ctypes.c_int, ctypes.POINTER(ZMat)]
    ret = lib.zblbicgstabl(ctypes.byref(Z), ctypes.byref(Bm), 2, 1e-10, 500, ctypes.byref(Xm))
    assert ret >= 0 and np.linalg.norm(B - A @ X) < 1e-8 * np.linalg.norm(B)
    assert "Block BiCGSTAB(2) has converged at:" in capfd.readouterr().out
    V = np.asfortranarray(zrand(np.random.default_rng(22), 200, 4))
    R = np.zeros((4, 4), np.complex128, order="F")
    V0 = V.copy(order="F")
    lib.zqreco.restype = ctypes.c_int
    assert lib.zqreco(200, 4, ctypes.c_void_p(V.ctypes.data), ctypes.c_void_p(R.ctypes.data)) == 0
    assert np.allclose(V @ R, V0, atol=1e-12) and np.allclose(V.conj().T @ V, np.eye(4), atol=1e-13)
    # zgeqrf + zungqr's convention (numpy's qr is that LAPACK pair): Q and R, signs
    # and the real diagonal included
    Qn, Rn = np.linalg.qr(V0)
    assert np.abs(R - Rn).max() <= 1e-12 * np.abs(Rn).max() and np.abs(V - Qn).max() <= 1e-12
    assert np.all(np.diag(R).imag == 0)


@pytest.mark.parametrize("m,s", [(5000, 3), (400, 20), (30, 30)])
def test_zqreco_lapack_convention(m, s):
    """src/myqr.c:41-76 (zgeqrf + zungqr) on the device, against numpy's LAPACK QR and
    the oracle's restatement, to 1e-12; a repeated column factors without failure."""
    lib = load()
    lib.zqreco.restype = ctypes.c_int
    V0 = np.asfortranarray(zrand(np.random.default_rng(23 + s), m, s))
    for rep in (False, True):
        if rep:
            V0[:, s - 1] = V0[:, 0]
        V = V0.copy(order="F")
        R = np.zeros((s, s), np.complex128, order="F")
        assert lib.zqreco(m, s, ctypes.c_void_p(V.ctypes.data), ctypes.c_void_p(R.ctypes.data)) == 0
        assert np.abs(V.conj().T @ V - np.eye(s)).max() < 1e-13
        assert np.abs(V @ R - V0).max() < 1e-12 * np.abs(V0).max() * 10
        if not rep:
            Qn, Rn = np.linalg.qr(V0)
            assert np.abs(R - Rn).max() <= 1e-12 * np.abs(Rn).max() and np.abs(V - Qn).max() <= 1e-12
            Qo, Ro = po.zqreco(V0)
            assert np.abs(R - Ro).max() <= 1e-12 * np.abs(Ro).max()
        else:
            assert abs(R[s - 1, s - 1]) < 1e-12 * np.abs(R).max()


@pytest.mark.parametrize("l", [1, 2, 4])
def test_zshbicgstabl_fused_shift_pass_bitwise(monkeypatch, l):
    """All per-shift work of an outer iteration in one pass (p_shift[k][1..l] in
    registers, the seed's r from snapshots) gives bitwise the iterates of the per-step
    kernels (BKH_ZSH_UNFUSED=1), after 2 outer iterations and at convergence."""
    n, rp, ci, v = zsystem(24, seed=8)
    b = zrand(np.random.default_rng(9), n)
    sig = np.array([0.02 + 0.1j, 0.1, 0.5 - 0.2j, 1.0 + 1.0j, 2.0, 0.3 + 0.7j, 0.7, 0.05j, 1.5 - 0.5j])
    out = {}
    for mode in ("fused", "unfused"):
        if mode == "unfused":
            monkeypatch.setenv("BKH_ZSH_UNFUSED", "1")
        out[mode] = [gpu_solve("shifted", n, rp, ci, v, b, l, sig, tol, mi) for tol, mi in ((1e-300, 2), (1e-10, 1000))]
    for (ia, xa, ta), (ib, xb, tb) in zip(out["fused"], out["unfused"]):
        assert ia.ret == ib.ret
        assert np.array_equal(xa, xb)
        assert np.array_equal(ta, tb)
    assert out["fused"][1][0].ret >= 0
