"""Training and evaluation pipeline of the learned FSAI preconditioner.

Fields are the variable-coefficient finite-volume diffusion matrices of
bksp_gen.h (kappa = exp(g), g a Gaussian random field). Training data per field:
its tril(A) and probe vectors z_p -- Rademacher vectors and normalised residuals
that unpreconditioned CG itself produces on the field (after a few and after more
iterations). The model (device MLP, BF16 matrix cores) is trained by the device
forward/backward/Adam kernels, one field per step.
"""
import numpy as np

from . import device as dv
from . import gen

# per-edge MLP 8 -> 64 -> 64 -> 1
NF, H = 8, 64
OFF = {"W1": 0, "b1": NF * H, "W2": NF * H + H, "b2": NF * H + H + H * H,
       "w3": NF * H + 2 * H + H * H, "b3": NF * H + 3 * H + H * H}
NPARAM = OFF["b3"] + 1


def init_params(seed=gen.SEEDS["cfg5_init"]):
    """Xavier-scaled tanh layers; a small output layer, so the untrained factor is
    close to Jacobi (h = 0 gives G = D^-1/2)."""
    z = gen.normal(seed, NPARAM)
    p = np.zeros(NPARAM, np.float32)
    p[OFF["W1"]:OFF["b1"]] = z[OFF["W1"]:OFF["b1"]] / np.sqrt(NF)
    p[OFF["W2"]:OFF["b2"]] = z[OFF["W2"]:OFF["b2"]] / np.sqrt(H)
    p[OFF["w3"]:OFF["b3"]] = 0.01 * z[OFF["w3"]:OFF["b3"]]
    return p


def field_system(seed, m, sigma=2.0, corr=8.0):
    """(n, lower CSR, full CSR) of the field with GRF seed `seed` on an m^3 grid."""
    kappa = gen.grf_kappa(seed, m, sigma, corr)
    n, rpl, cil, vl = gen.fv_diffusion(m, kappa, lower=True)
    return n, (rpl, cil, vl)


def probes(n, lower, seed, n_rademacher=2, residual_iters=(8, 40)):
    """Rademacher vectors and normalised CG residuals r_k = b - A x_k (k iterations of
    unpreconditioned CG from x0 = 0, b ~ N(0,1)), scaled to |z|^2 = n."""
    rpl, cil, vl = lower
    Z = []
    u = gen.uniform(seed, n * n_rademacher).reshape(n_rademacher, n)
    Z += list(np.where(u < 0.5, -1.0, 1.0))
    if residual_iters:
        M = dv.Matrix(n, rpl, cil, vl, symmetric=True)
        b = gen.normal(seed + 1, n)
        db = dv.DeviceArray.from_numpy(b)
        dx, dy = dv.DeviceArray((n,)), dv.DeviceArray((n,))
        for k in residual_iters:
            dx.upload(np.zeros(n))
            dv.cg(M, db, dx, 1e-30, k)
            M.spmv(dx, dy)
            r = b - dy.numpy()
            Z.append(r * np.sqrt(n) / np.linalg.norm(r))
        M.destroy()
    return np.ascontiguousarray(np.array(Z))


def make_training_fields(seeds, m):
    fields = []
    for s in seeds:
        n, lower = field_system(s, m)
        fields.append(dv.FsaiField(n, *lower, Z=probes(n, lower, 10 * s + 1)))
    return fields


def train(model, fields, epochs, lr=3e-3, lr_final=3e-4, log_every=0):
    """Adam over the fields (one field per step), learning rate decayed
    geometrically from lr to lr_final. Returns the per-epoch mean losses."""
    steps = epochs * len(fields)
    hist, k = [], 0
    for ep in range(epochs):
        tot = 0.0
        for f in fields:
            rate = lr * (lr_final / lr) ** (k / max(1, steps - 1))
            L = f.train_step(model, rate, want_loss=True)
            tot += L
            k += 1
        hist.append(tot / len(fields))
        if log_every and (ep + 1) % log_every == 0:
            print(f"[fsai] epoch {ep + 1}: mean loss {hist[-1]:.5f}", flush=True)
    return hist


def solve_compare(model, seed, m, tol=1e-8, max_iter=200000, b_seed=gen.SEEDS["cfg5_rhs"], plain=True,
                  system=None, untrained=None):
    """Unpreconditioned CG, Jacobi PCG and learned-FSAI PCG on one held-out field
    (`system`: the field's (n, lower CSR) if already generated; `untrained`: a second
    model, e.g. the initial weights, whose factor is solved with as well)."""
    n, lower = system if system is not None else field_system(seed, m)
    F = dv.FsaiField(n, *lower)
    b = gen.normal(b_seed, n)
    db, dx = dv.DeviceArray.from_numpy(b), dv.DeviceArray((n,))
    zeros = np.zeros(n)
    out = {}

    def run(name, fn):
        dx.upload(zeros)
        info = fn()
        out[name] = {"iters": info.ret, "ms": info.time_ms, "true_res": info.true_res}

    build = {}
    if untrained is not None:
        F.factor(untrained)
        run("untrained", lambda: F.pcg(db, dx, tol, max_iter))
    F.factor(model)
    build["ms"] = F.bench_forward(model, reps=3)
    run("jacobi", lambda: F.pcg_plain(1, db, dx, tol, max_iter))
    run("learned", lambda: F.pcg(db, dx, tol, max_iter))
    out["learned"]["factor_build_ms"] = build["ms"]
    out["learned"]["edges"] = F.nnzl
    if plain:
        run("plain_cg", lambda: F.pcg_plain(0, db, dx, tol, max_iter))
    F.destroy()
    return out
