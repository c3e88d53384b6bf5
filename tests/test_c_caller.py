"""A C program compiled against include/bksp.h with gcc and linked to
libbksp_amd.so, as a program written for the reference library (its driver is
tests/test.c:51-310) links -lbksp: tests/c_caller/caller.c.

CPU: the program compiles with -Wall -Wextra -Werror and links with no undefined
symbol. GPU: it runs the driver's scenarios (bfwb398 and pde2961, b = e1, 16
right-hand sides, tol 1e-14) through the drop-in, and every result is compared
with the oracle run on the same inputs: the break index within the fixture bar
(+-max(2, 10%)), x within 1e-8 relative of the oracle's iterate at the device's
index, the exported SpMV symbols bitwise and sp_norm_f to 1e-14."""
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT, oracle_at_index, read_mtx
from oracle import pyoracle as po

CALLER_DIR = os.path.join(ROOT, "tests", "c_caller")
LIB = os.path.join(ROOT, "bksp_amd", "libbksp_amd.so")
SYM = os.path.join(GOLDEN, "testmat", "bfwb398.mtx")
GEN = os.path.join(GOLDEN, "testmat", "pde2961.mtx")


def build(out):
    r = subprocess.run(["make", "-s", "-C", CALLER_DIR, f"OUT={out}"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    return out


def test_caller_compiles_and_links(tmp_path):
    if not os.path.exists(LIB):
        pytest.skip("libbksp_amd.so is not built (run __graft_entry__.build())")
    exe = build(str(tmp_path / "bksp_caller"))
    assert os.access(exe, os.X_OK)


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def fixture_index(a, b):
    return abs(a - b) <= max(2, int(0.1 * max(a, b)))


@pytest.fixture(scope="module")
def caller_run(tmp_path_factory):
    exe = os.path.join(CALLER_DIR, "bksp_caller")
    if not os.path.exists(exe):
        exe = build(str(tmp_path_factory.mktemp("bin") / "bksp_caller"))
    out = tmp_path_factory.mktemp("caller")
    print("running the C caller on the bfwb398 / pde2961 scenarios", flush=True)
    r = subprocess.run([exe, SYM, GEN, str(out)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    res = {}
    for line in open(out / "results.txt"):
        f = line.split()
        if f[0].startswith("fnorm"):
            res[f[0]] = float(f[1])
        else:
            name, ret, n, cols = f[0], int(f[1]), int(f[2]), int(f[3])
            x = np.fromfile(out / f"{name}.bin", dtype=np.float64).reshape(cols, n).T
            res[name] = (ret, x[:, 0] if cols == 1 else x)
    n = read_mtx(GEN)[0]
    res["B"] = np.fromfile(out / "B.bin", dtype=np.float64).reshape(16, n).T.copy()
    res["spmv_x"] = np.fromfile(out / "spmv_x.bin", dtype=np.float64)
    return res


@pytest.mark.gpu
def test_caller_symmetric_scenarios(caller_run, bfwb398):
    n, rp, ci, v, sym = bfwb398
    assert sym
    # the device sums the squares as a tree: equal to the oracle's sequential sum to rounding
    assert abs(caller_run["fnorm_sym"] - po.sp_norm_f(n, rp, ci, v, True)) <= 1e-14 * caller_run["fnorm_sym"]
    b = np.zeros(n); b[0] = 1.0
    for name, fn in (("cg", po.cg), ("cr", po.cr)):
        ret, x = caller_run[name]
        cret, _ = fn(n, rp, ci, v, b, 1e-14, 500)
        assert ret >= 0 and cret >= 0 and fixture_index(ret, cret), (name, ret, cret)
        xk = oracle_at_index(lambda t, m: fn(n, rp, ci, v, b, t, m), ret)
        assert rel(x, xk) < 1e-8, (name, rel(x, xk))
    ret, X = caller_run["blcg"]
    B = np.zeros((n, 16)); B[np.arange(16), np.arange(16)] = 1.0
    cret, _ = po.blcg(n, rp, ci, v, B, 1e-14, 500)
    assert ret >= 0 and cret >= 0 and fixture_index(ret, cret), (ret, cret)
    Xk = oracle_at_index(lambda t, m: po.blcg(n, rp, ci, v, B, t, m), ret)
    assert rel(X, Xk) < 1e-8, rel(X, Xk)


@pytest.mark.gpu
def test_caller_general_scenarios(caller_run, pde2961):
    n, rp, ci, v, sym = pde2961
    assert not sym
    assert abs(caller_run["fnorm_gen"] - po.sp_norm_f(n, rp, ci, v, False)) <= 1e-14 * caller_run["fnorm_gen"]
    b = np.zeros(n); b[0] = 1.0
    B = caller_run["B"]
    runs = {
        "bicg": lambda t, m: po.bicg(n, rp, ci, v, b, t, m),
        "bicgstab": lambda t, m: po.bicgstab(n, rp, ci, v, b, t, m),
        "bicgstabl": lambda t, m: po.bicgstabl(n, rp, ci, v, b, 4, t, m),
        "blbicgstab": lambda t, m: po.blbicgstab(n, rp, ci, v, B, t, m)[:2],
        "blbicgstabl": lambda t, m: po.blbicgstabl(n, rp, ci, v, B, 6, t, m)[:2],
        "shbicg": lambda t, m: po.shbicg(n, rp, ci, v, b, [0.0, 0.2, 0.4, 0.8], t, m),
        "shbicgstabl": lambda t, m: po.shbicgstabl(n, rp, ci, v, b, 1, [0.0, 0.2, 0.4, 0.8], t, m),
    }
    maxit = {"shbicgstabl": 200}
    for name, call in runs.items():
        ret, x = caller_run[name]
        cret = call(1e-14, maxit.get(name, 500))[0]
        assert (ret >= 0) == (cret >= 0), (name, ret, cret)
        if ret >= 0:
            assert fixture_index(ret, cret), (name, ret, cret)
        k = ret if ret >= 0 else maxit.get(name, 500) - 1
        xk = oracle_at_index(call, k)
        if name.startswith("sh"):
            xk = xk.T                       # oracle: (shifts + 1) x n; caller: n x (shifts + 1)
        err = rel(x, xk)
        print(f"{name}: device {ret}, oracle {cret}, |x - x_oracle| at index {k} = {err:.2e}")
        assert err < 1e-8, (name, err)
    x = caller_run["spmv_x"]
    assert np.array_equal(caller_run["spmv_n"][1], po.csrgemv("n", n, rp, ci, v, x))
    assert np.array_equal(caller_run["spmv_t"][1], po.csrgemv("t", n, rp, ci, v, x))
