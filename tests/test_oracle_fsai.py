"""The oracle's learned-FSAI model (oracle/orc_fsai.c) pinned on the CPU: the
parametrisation's Jacobi limit, the Kaporin inequality and the classical FSAI
optimum of the loss, the probe estimate of the trace, and the backward pass
against central finite differences."""
import numpy as np
import pytest

from bksp_amd import gen
from oracle import pyoracle as po


def field(m, seed):
    kappa = gen.grf_kappa(seed, m, 2.0, 2.0)
    n, rpl, cil, vl = gen.fv_diffusion(m, kappa, lower=True)
    _, rp, ci, v = gen.fv_diffusion(m, kappa, lower=False)
    return n, rp, ci, v, rpl, cil, vl


def dense(n, rp, ci, v):
    A = np.zeros((n, n))
    for i in range(n):
        A[i, ci[rp[i]:rp[i + 1]]] += v[rp[i]:rp[i + 1]]
    return A


def lower_dense(n, rpl, cil, g):
    G = np.zeros((n, n))
    for i in range(n):
        G[i, cil[rpl[i]:rpl[i + 1]]] = g[rpl[i]:rpl[i + 1]]
    return G


def rademacher(P, n, seed):
    return np.where(gen.uniform(seed, P * n) < 0.5, -1.0, 1.0).reshape(P, n)


def zero_params():
    return np.zeros(po.fsai_nparam(), np.float32)


def random_params(seed, scale=0.2):
    return (gen.normal(seed, po.fsai_nparam()) * scale).astype(np.float32)


def exact_loss(A, G):
    n = A.shape[0]
    B = G @ A @ G.T
    return np.log(np.trace(B) / n) - 2.0 * np.sum(np.log(np.diag(G))) / n


def test_zero_network_is_jacobi():
    n, rp, ci, v, rpl, cil, vl = field(4, 1)
    rows = po.fsai_rows(n, rp, ci, v)
    h = po.fsai_forward(n, rpl, cil, vl, rows, zero_params(), False)
    assert np.all(h == 0.0)
    g = po.fsai_gvals(n, rpl, cil, rows[0], h)
    A = dense(n, rp, ci, v)
    G = lower_dense(n, rpl, cil, g)
    assert np.allclose(G, np.diag(1.0 / np.sqrt(np.diag(A))), rtol=1e-14, atol=0)


def test_loss_is_kaporin_number_and_fsai_is_its_minimum():
    """With exact traces L + log(det A)/n = log K(G A G^T) >= 0 (AM-GM), and the
    classical FSAI factor (rows from the small dense systems on tril(A)'s pattern)
    minimises it over the pattern: it beats Jacobi and random perturbations."""
    n, rp, ci, v, rpl, cil, vl = field(4, 2)
    A = dense(n, rp, ci, v)
    logdet = np.linalg.slogdet(A)[1]
    rows = po.fsai_rows(n, rp, ci, v)
    g = po.fsai_gvals(n, rpl, cil, rows[0], po.fsai_forward(n, rpl, cil, vl, rows, random_params(3), False))
    G = lower_dense(n, rpl, cil, g)
    assert exact_loss(A, G) + logdet / n >= -1e-12
    # classical FSAI: for row i with pattern J (i last), A[J,J] y = e_last, g = y / sqrt(y_last)
    Gf = np.zeros((n, n))
    for i in range(n):
        J = cil[rpl[i]:rpl[i + 1]]
        y = np.linalg.solve(A[np.ix_(J, J)], (J == i).astype(float))
        Gf[i, J] = y / np.sqrt(y[J == i][0])
    Lf = exact_loss(A, Gf)
    Lj = exact_loss(A, np.diag(1.0 / np.sqrt(np.diag(A))))
    assert Lf < Lj
    rng = np.random.default_rng(0)
    for _ in range(5):
        Gp = Gf + 1e-3 * (Gf != 0) * rng.standard_normal((n, n))
        assert exact_loss(A, Gp) >= Lf - 1e-12


def test_probe_trace_estimate():
    n, rp, ci, v, rpl, cil, vl = field(5, 4)
    A = dense(n, rp, ci, v)
    rows = po.fsai_rows(n, rp, ci, v)
    g = po.fsai_gvals(n, rpl, cil, rows[0], po.fsai_forward(n, rpl, cil, vl, rows, random_params(5), False))
    G = lower_dense(n, rpl, cil, g)
    Z = rademacher(400, n, 9)
    L, _ = po.fsai_loss_dg(n, rp, ci, v, rpl, cil, g, Z)
    assert abs(L - exact_loss(A, G)) < 0.03


@pytest.mark.parametrize("seed", [11, 12])
def test_backward_matches_finite_differences(seed):
    n, rp, ci, v, rpl, cil, vl = field(4, seed)
    Z = rademacher(3, n, seed + 100)
    p = random_params(seed + 7)
    L0, grad, _ = po.fsai_loss_grad(n, rp, ci, v, rpl, cil, vl, p, Z, bf16=False)
    rng = np.random.default_rng(seed)
    offsets = {"W1": (0, 512), "b1": (512, 576), "W2": (576, 4672), "b2": (4672, 4736),
               "w3": (4736, 4800), "b3": (4800, 4801)}
    for name, (a, b) in offsets.items():
        for k in rng.choice(np.arange(a, b), size=min(6, b - a), replace=False):
            pp, pm = p.copy(), p.copy()
            pp[k] += np.float32(2e-3)
            pm[k] -= np.float32(2e-3)
            step = float(pp[k]) - float(pm[k])
            Lp = po.fsai_loss_grad(n, rp, ci, v, rpl, cil, vl, pp, Z)[0]
            Lm = po.fsai_loss_grad(n, rp, ci, v, rpl, cil, vl, pm, Z)[0]
            fd = (Lp - Lm) / step
            assert abs(fd - grad[k]) <= 2e-4 * max(1e-3, abs(grad[k])) + 1e-7, (name, k, fd, grad[k])


def test_bf16_mode_stays_close_to_double():
    n, rp, ci, v, rpl, cil, vl = field(4, 21)
    Z = rademacher(3, n, 22)
    p = random_params(23)
    L64, g64, _ = po.fsai_loss_grad(n, rp, ci, v, rpl, cil, vl, p, Z, bf16=False)
    L16, g16, _ = po.fsai_loss_grad(n, rp, ci, v, rpl, cil, vl, p, Z, bf16=True)
    assert abs(L16 - L64) < 1e-2 * abs(L64) + 1e-3
    assert np.linalg.norm(g16 - g64) < 5e-2 * np.linalg.norm(g64)
