"""Seeded benchmark systems, built by the C generators of include/bksp_gen.h.

Each function returns ``(n, row_ptr, col, val)`` as NumPy arrays (0-based CSR,
int32 indices, ascending columns). The same arrays feed the GPU path and the
CPU oracle, so both see identical bytes.
"""
import ctypes

import numpy as np

from ._lib import CsrHost, load

SEEDS = {"cfg1_xtrue": 101, "cfg2_rhs": 202, "cfg3_rhs": 303, "cfg4_struct": 404,
         "cfg4_rhs": 405, "cfg5_train0": 5000, "cfg5_test0": 6000, "cfg5_rhs": 505,
         "cfg5_init": 507}


def _take(m):
    n, nnz = m.n, m.nnz
    rp = np.ctypeslib.as_array(m.row_ptr, shape=(n + 1,)).copy()
    ci = np.ctypeslib.as_array(m.col, shape=(nnz,)).copy() if nnz else np.zeros(0, np.int32)
    v = np.ctypeslib.as_array(m.val, shape=(nnz,)).copy() if nnz else np.zeros(0)
    load().bkh_csr_host_free(ctypes.byref(m))
    return n, rp, ci, v


def _call(fn, *args):
    m = CsrHost()
    if fn(*args, ctypes.byref(m)) != 0:
        raise RuntimeError(f"generator {fn.__name__} failed")
    return _take(m)


def convdiff2d(m=64, vx=10.0, vy=20.0):
    return _call(load().bkh_gen_convdiff2d, m, vx, vy)


def poisson3d(m, lower=True):
    return _call(load().bkh_gen_poisson3d, m, int(lower))


def convdiff3d(m, px, py, pz):
    return _call(load().bkh_gen_convdiff3d, m, px, py, pz)


def random_rows(seed, n):
    return _call(load().bkh_gen_random_rows, seed, n)


def batch_sizes(seed, count, lo2=10.0, hi2=17.0):
    out = np.zeros(count, np.int32)
    load().bkh_gen_batch_sizes(seed, count, lo2, hi2, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    return out


def uniform(seed, n, lo=0.0, hi=1.0):
    out = np.empty(int(n))
    load().bkh_gen_uniform(seed, int(n), lo, hi, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return out


def normal(seed, n):
    out = np.empty(int(n))
    load().bkh_gen_normal(seed, int(n), out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return out


def grf_kappa(seed, m, sigma=2.0, corr=8.0):
    out = np.empty(m * m * m)
    if load().bkh_gen_grf(seed, m, sigma, corr, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))):
        raise RuntimeError("bkh_gen_grf failed")
    return out


def fv_diffusion(m, kappa, lower=True):
    k = np.ascontiguousarray(kappa, dtype=np.float64)
    return _call(load().bkh_gen_fv_diffusion, m, k.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                 int(lower))


def cfg3_peclet(pmax=0.4, v=(1.0, 2.0, 3.0)):
    """Cell Peclet numbers p_i = |v_i| h / (2 nu) scaled so that max p_i = pmax."""
    s = pmax / max(v)
    return tuple(s * vi for vi in v)


def cfg4_system_seed(s, base=SEEDS["cfg4_struct"]):
    """Seed of system s of the uneven batch (structure and values)."""
    return base * 65536 + s


def cfg4_batch(count=512, seed=SEEDS["cfg4_struct"], lo2=10.0, hi2=17.0):
    """The uneven batch: sizes n_s = round(2^U[lo2, hi2]) from `seed`, system s from
    random_rows(cfg4_system_seed(s), n_s). Returns a list of (n, row_ptr, col, val)."""
    sizes = batch_sizes(seed, count, lo2, hi2)
    return [random_rows(cfg4_system_seed(q, seed), int(n)) for q, n in enumerate(sizes)]


def zconvdiff3d(m, delta=0.05, pmax=0.4):
    """Complex test operator: the m^3 convection-diffusion matrix of cfg3_peclet
    with an imaginary diagonal shift i*delta*a_ii (a complex nonsymmetric system of
    the shifted-Helmholtz kind). Returns (n, row_ptr, col, complex val)."""
    px, py, pz = cfg3_peclet(pmax)
    n, rp, ci, v = convdiff3d(m, px, py, pz)
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(rp))
    z = v.astype(np.complex128)
    d = ci == rows
    z[d] += 1j * delta * v[d]
    return n, rp, ci, z
