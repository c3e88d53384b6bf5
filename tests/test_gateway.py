"""The MATLAB-gateway front-end (bksp_amd/gateway.py): argument checks with the
gateways' identifiers and messages (CPU, no device call), and the solves through
the drop-in library (GPU) against the oracle."""
import numpy as np
import pytest
import scipy.sparse as sp

from bksp_amd import gateway as gw
from conftest import zsystem


def small():
    from bksp_amd import gen
    n, rp, ci, v = gen.convdiff2d(16, 10.0, 20.0)
    return sp.csr_matrix((v, ci, rp), shape=(n, n))


@pytest.mark.parametrize("fn,args,ident,msg", [
    (gw.bicg, (1, 2, 3), "BKSP:bicg", "Four inputs required."),
    (gw.bicgstabl, (1, 2, 3, 4), "BKSP:bicgstabl", "Five inputs required."),
    (gw.blbicgstabl, (1, 2, 3, 4), "BKSP:blbicgstabl", "Five inputs required."),
    (gw.shbicgstabl, (1, 2, 3, 4, 5), "BKSP:shbicgstabl", "Six inputs required."),
])
def test_input_count(fn, args, ident, msg):
    with pytest.raises(gw.GatewayError) as e:
        fn(*args)
    assert e.value.ident == ident and e.value.msg == msg


def test_argument_checks():
    A = small()
    n = A.shape[0]
    cases = [
        (lambda: gw.bicg(A.toarray(), np.ones(n), 1e-8, 10), "BKSP:invalidInputSparisty",
         "Input argument must be a sparse square matrix."),
        (lambda: gw.bicg(sp.csr_matrix(np.ones((3, 4))), np.ones(3), 1e-8, 10), "BKSP:invalidInputSparisty",
         "Input argument must be a sparse square matrix."),
        (lambda: gw.bicgstabl(A, np.ones((n, 2)), 2, 1e-8, 10), "BKSP:invalidInputVector",
         "Input right hand side must be a vector."),
        (lambda: gw.bicgstabl(A, np.ones(n + 1), 2, 1e-8, 10), "BKSP:invalidInputVector",
         "Input right hand side must have the same size of the coefficient matrix."),
        (lambda: gw.bicgstabl(A, np.ones(n), [2, 3], 1e-8, 10), "BKSP:bicgstabl",
         "Input parameter ell must be an integer."),
        (lambda: gw.blbicgstabl(A, np.ones((n, 2)), 2, [1e-8, 1e-9], 10), "BKSP:blbicgstabl",
         "Input tolerance must be a scalar."),
        (lambda: gw.shbicgstabl(A, np.ones(n), 2, [0.1], 1e-8, [10, 20]), "BKSP:blbicgstabl",
         "Input max iteration must be an integer."),
    ]
    for f, ident, msg in cases:
        with pytest.raises(gw.GatewayError) as e:
            f()
        assert (e.value.ident, e.value.msg) == (ident, msg)


@pytest.mark.gpu
def test_gateway_solves(capfd):
    from oracle import pyoracle as po
    A = small()
    n = A.shape[0]
    b = np.random.default_rng(1).standard_normal(n)
    x = gw.bicgstabl(A, b, 2, 1e-10, 500)
    assert np.linalg.norm(b - A @ x) < 1e-9 * np.linalg.norm(b)
    assert "The BiCGSTAB(l) Method converged at iteration" in capfd.readouterr().out
    X = gw.shbicgstabl(A, b, 2, [0.1, 1.0], 1e-10, 500)
    for k, s in enumerate([0.0, 0.1, 1.0]):
        assert np.linalg.norm(b - A @ X[:, k] - s * X[:, k]) < 1e-8 * np.linalg.norm(b)
    n2, rp, ci, v = zsystem(16, seed=2)
    Z = sp.csr_matrix((v, ci, rp), shape=(n2, n2))
    zb = np.random.default_rng(3).standard_normal(n2) + 1j
    zx = gw.bicg(Z, zb, 1e-10, 1000)
    ret, cx = po.zbicg(n2, rp, ci, v, zb, 1e-10, 1000)
    assert np.linalg.norm(zb - Z @ zx) < 1e-9 * np.linalg.norm(zb) and np.linalg.norm(zx - cx) < 1e-7 * np.linalg.norm(cx)
    ZB = np.random.default_rng(4).standard_normal((n2, 2)) + 0.5j
    ZX = gw.blbicgstabl(Z, ZB, 2, 1e-10, 500)
    assert np.linalg.norm(ZB - Z @ ZX) < 1e-8 * np.linalg.norm(ZB)
    capfd.readouterr()
