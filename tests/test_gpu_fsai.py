"""Learned FSAI on the device against the oracle (oracle/orc_fsai.c), and what the
trained network gains over the method without it (h = 0, i.e. Jacobi PCG)."""
import numpy as np
import pytest

from bksp_amd import device as dv
from bksp_amd import fsai
from bksp_amd import gen
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def small_field(seed, m=10):
    n, (rpl, cil, vl) = fsai.field_system(seed, m)
    _, rp, ci, v = gen.fv_diffusion(m, gen.grf_kappa(seed, m), lower=False)
    return n, rp, ci, v, rpl, cil, vl


def rademacher(P, n, seed):
    return np.where(gen.uniform(seed, P * n) < 0.5, -1.0, 1.0).reshape(P, n)


def test_factor_matches_oracle_and_zero_network_is_jacobi():
    n, rp, ci, v, rpl, cil, vl = small_field(31)
    F = dv.FsaiField(n, rpl, cil, vl)
    p = fsai.init_params(9) * 4          # a non-trivial network
    M = dv.FsaiModel(p)
    g = F.factor(M, want=True)
    rows = po.fsai_rows(n, rp, ci, v)
    gref = po.fsai_gvals(n, rpl, cil, rows[0], po.fsai_forward(n, rpl, cil, vl, rows, p, True))
    err = np.abs(g - gref) / np.abs(gref).max()
    print(f"factor: max err {err.max():.2e}, median {np.median(err):.2e}")
    assert err.max() < 2e-3 and np.median(err) < 1e-5
    M0 = dv.FsaiModel(np.zeros(fsai.NPARAM, np.float32))
    g0 = F.factor(M0, want=True)
    diag = rpl[1:] - 1                   # the diagonal is the last entry of each lower row
    assert np.allclose(g0[diag], 1.0 / rows[0], rtol=1e-6)
    off = np.ones(len(g0), bool); off[diag] = False
    assert np.all(g0[off] == 0.0)


@pytest.mark.parametrize("seed", [41, 42])
def test_loss_and_gradient_match_oracle(seed):
    """Device forward/backward kernels (BF16 matrix cores, FP32 accumulation) against
    the oracle's BF16-emulating restatement; the oracle's own backward is pinned by
    finite differences in tests/test_oracle_fsai.py."""
    n, rp, ci, v, rpl, cil, vl = small_field(seed)
    Z = rademacher(3, n, seed + 5)
    p = fsai.init_params(seed) * 3
    F = dv.FsaiField(n, rpl, cil, vl, Z=Z)
    M = dv.FsaiModel(p)
    L, g = F.loss_grad(M)
    Lr, gr, _ = po.fsai_loss_grad(n, rp, ci, v, rpl, cil, vl, p, Z, bf16=True)
    rel = np.linalg.norm(g - gr) / np.linalg.norm(gr)
    print(f"loss gpu {L:.8f} oracle {Lr:.8f}; grad rel err {rel:.2e}")
    assert abs(L - Lr) < 1e-5 * max(1.0, abs(Lr))
    assert rel < 1e-2
    # every parameter block is exercised
    for name, a in fsai.OFF.items():
        b = a + {"W1": 512, "b1": 64, "W2": 4096, "b2": 64, "w3": 64, "b3": 1}[name]
        assert np.linalg.norm(g[a:b] - gr[a:b]) <= 2e-2 * np.linalg.norm(gr[a:b]) + 1e-9, name


def test_pcg_none_is_cg_and_jacobi_converges():
    n, rp, ci, v, rpl, cil, vl = small_field(51, 16)
    F = dv.FsaiField(n, rpl, cil, vl)
    b = gen.normal(505, n)
    db, dx = dv.DeviceArray.from_numpy(b), dv.DeviceArray.from_numpy(np.zeros(n))
    i0 = F.pcg_plain(0, db, dx, 1e-8, 100000)
    x0 = dx.numpy()
    cret, cx = po.cg(n, rpl, cil, vl, b, 1e-8, 100000)
    assert i0.ret >= 0 and abs(i0.ret - cret) <= max(1, int(0.05 * cret)), (i0.ret, cret)
    assert i0.true_res < 1.5e-8
    dx.upload(np.zeros(n))
    i1 = F.pcg_plain(1, db, dx, 1e-8, 100000)
    assert i1.ret >= 0 and i1.true_res < 1.5e-8 and i1.ret < i0.ret
    assert np.linalg.norm(dx.numpy() - x0) / np.linalg.norm(x0) < 1e-5


def test_training_reduces_loss_and_beats_jacobi_on_held_out_field():
    train = fsai.make_training_fields(range(5000, 5008), 16)
    M = dv.FsaiModel(fsai.init_params())
    hist = fsai.train(M, train, epochs=25, lr=5e-3, lr_final=5e-4)
    print("loss per epoch", [round(h, 4) for h in hist[::5]], hist[-1])
    assert hist[-1] < hist[0] - 0.02
    res = fsai.solve_compare(M, 6000, 24, plain=False)
    print(res)
    assert res["learned"]["true_res"] < 1e-8 and res["jacobi"]["true_res"] < 1e-8
    assert res["learned"]["iters"] < 0.9 * res["jacobi"]["iters"]


@pytest.mark.parametrize("kind", ["none", "jacobi", "fsai"])
def test_fused_pcg_matches_unfused(kind):
    """The fused PCG (path 2: two or three kernels per iteration, SELL matrices; FSAI
    keeps z and y in FP32) against the unfused loop (path 1) on the same field:
    iteration indices within 2 % and the same solution to the tolerance's accuracy;
    both meet tol in the FP64 true residual."""
    n, rp, ci, v, rpl, cil, vl = small_field(61, 20)
    F = dv.FsaiField(n, rpl, cil, vl)
    b = gen.normal(505, n)
    db = dv.DeviceArray.from_numpy(b)
    out = []
    for path in (2, 1):
        dx = dv.DeviceArray.from_numpy(np.zeros(n))
        if kind == "fsai":
            M = dv.FsaiModel(fsai.init_params(3) * 2)
            F.factor(M)
            info = F.pcg(db, dx, 1e-9, 100000, path=path)
        else:
            info = F.pcg_plain(0 if kind == "none" else 1, db, dx, 1e-9, 100000, path=path)
        out.append((info, dx.numpy()))
    (i0, x0), (i1, x1) = out
    print(f"{kind}: fused {i0.ret} iters {i0.time_ms:.2f} ms, unfused {i1.ret} iters {i1.time_ms:.2f} ms")
    assert i0.ret >= 0 and i1.ret >= 0
    assert abs(i0.ret - i1.ret) <= max(1, int(0.02 * i1.ret)), (i0.ret, i1.ret)
    assert i0.true_res < 1.5e-9 and i1.true_res < 1.5e-9
    assert np.linalg.norm(x0 - x1) / np.linalg.norm(x1) < 1e-6
