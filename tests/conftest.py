import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libbksp_amd.so on the device)")


def read_mtx(path):
    """Matrix Market coordinate file -> 0-based CSR with the reference's in-row
    order (src/read_matrix.c:122-173: a stable counting sort by row, so each row
    keeps file order). Returns (n, row_ptr, col, val, symmetric)."""
    with open(path) as f:
        banner = f.readline().split()
        line = f.readline()
        while line.startswith("%"):
            line = f.readline()
        M, N, nz = (int(t) for t in line.split())
        data = np.loadtxt(f, ndmin=2)
    r = data[:, 0].astype(np.int64) - 1
    c = data[:, 1].astype(np.int32) - 1
    v = data[:, 2].astype(np.float64)
    order = np.argsort(r, kind="stable")
    rp = np.zeros(M + 1, np.int32)
    np.add.at(rp, r + 1, 1)
    rp = np.cumsum(rp).astype(np.int32)
    sym = banner[4].lower() == "symmetric"
    return M, rp, c[order], v[order], sym


def oracle_at_index(call, k):
    """The oracle's iterate after loop index k, where the device broke.

    ``call(tol, max_iter)`` runs an oracle solver and returns (ret, x, ...). Every
    reference loop runs ``for (i = start; i < max_iter; i++)`` and tests
    ``error < tol`` after the iteration's update (cg.c:33-60, cr.c:37-68,
    bicg.c:35-73, bicgstabl.c:53-118, blcg.c:66-106, blbicgstabl.c:96-212,
    shbicgstabl.c:83-247), so with tol 0 the break never fires and max_iter = k + 1
    performs exactly the iterations start..k that the device performed before
    breaking at k. Returns the oracle's x at that index."""
    assert k >= 0, k
    return call(0.0, k + 1)[1]


def count_index_mismatch(ret, cret):
    """Record (in the test log) a device/oracle break-index difference; the x check
    runs at the device's index either way."""
    if ret != cret:
        print(f"break index: device {ret}, oracle {cret} (x compared at the device's index)")


@pytest.fixture(scope="session")
def bfwb398():
    return read_mtx(os.path.join(GOLDEN, "testmat", "bfwb398.mtx"))


@pytest.fixture(scope="session")
def pde2961():
    return read_mtx(os.path.join(GOLDEN, "testmat", "pde2961.mtx"))


def expand_lower(n, rp, ci, v):
    """Full matrix of a stored triangle as scipy CSR (independent of the oracle)."""
    import scipy.sparse as sp
    rows = np.repeat(np.arange(n), np.diff(rp))
    A = sp.coo_matrix((v, (rows, ci)), shape=(n, n)).tocsr()
    off = sp.coo_matrix((v[rows != ci], (ci[rows != ci], rows[rows != ci])), shape=(n, n)).tocsr()
    return (A + off).tocsr()


def to_scipy(n, rp, ci, v):
    import scipy.sparse as sp
    return sp.csr_matrix((v, ci, rp), shape=(n, n))


def zsystem(m, seed=0, delta=0.3, offim=0.2, vx=10.0, vy=20.0):
    """A complex nonsymmetric test system: the m x m 2D convection-diffusion
    operator with an imaginary diagonal shift i*delta and seeded imaginary parts
    (|.| <= offim times the real part) on its off-diagonal entries.
    Returns (n, row_ptr, col, complex val)."""
    from bksp_amd import gen
    n, rp, ci, v = gen.convdiff2d(m, vx, vy)
    rng = np.random.default_rng(seed)
    rows = np.repeat(np.arange(n), np.diff(rp))
    z = v.astype(np.complex128)
    off = ci != rows
    z[off] += 1j * offim * v[off] * rng.uniform(-1.0, 1.0, int(off.sum()))
    z[~off] += 1j * delta * np.abs(v[~off])
    return n, rp.astype(np.int32), ci.astype(np.int32), z


def zcsr_dense(n, rp, ci, v):
    import scipy.sparse as sp
    return sp.csr_matrix((v, ci, rp), shape=(n, n))
