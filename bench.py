"""Benchmark: CG to a 1e-8 relative residual on the 128^3 7-point Poisson system
(2,097,152 rows, 14,581,760 expanded nonzeros, passed as its 8,339,456-entry
lower triangle), RHS N(0,1) seed 202, x0 = 0, on one MI355X.

One step = one complete solve from x0 = 0 to tol = 1e-8 (plain FP64 GPU path,
bkh_cg on a matrix and vectors already resident in HBM). The bksp_dcg drop-in
(host arrays: upload, layout, solve, download) is timed beside it and is not
`value`. Prints one JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "time to 1e-8 true rel. residual vs plain FP64 GPU; SpMV/SpMM HBM GB/s"
HBM_PEAK_GBS = 8000.0


def spmv_bytes(n, nnz, val_bytes=8, vec_bytes=8):
    """Algorithmic bytes of one CSR SpMV: values + col indices + row_ptr + x + y."""
    return val_bytes * nnz + 4 * nnz + 4 * (n + 1) + vec_bytes * n + vec_bytes * n


def load_traffic(name):
    p = os.path.join(ROOT, "profiles", name)
    if os.path.exists(p):
        with open(p) as f:
            return json.load(f)
    return None


def _sha256(*arrays):
    import hashlib
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _host_info():
    """CPU model and the cores a threaded CPU run may use on this host."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    threads = int(os.environ.get("OMP_NUM_THREADS", avail))
    return model, min(threads, avail), avail


def _timed(fn, reps=1):
    from bksp_amd import load
    load().bkh_sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        r = fn()
    load().bkh_sync()
    return r, (time.perf_counter() - t0) * 1e3 / reps


def bench_cfg1(dv, gen, po):
    """BiCGSTAB(2), 64x64 convection-diffusion (n = 4096), tol 1e-10: latency case."""
    n, rp, ci, v = gen.convdiff2d(64, 10.0, 20.0)
    xt = gen.uniform(101, n, -1.0, 1.0)
    import scipy.sparse as sp
    b = sp.csr_matrix((v, ci, rp), shape=(n, n)) @ xt
    M = dv.Matrix(n, rp, ci, v)
    db, dx = dv.DeviceArray.from_numpy(b), dv.DeviceArray.from_numpy(np.zeros(n))
    def run(**kw):
        dx.zero()
        return dv.bicgstabl(M, db, 2, dx, 1e-10, 1000, **kw)

    run()
    info, ms = _timed(run, 10)
    info_g, ms_g = _timed(lambda: run(path=1), 5)
    cpu = None
    if po is not None:                       # CPU baseline leg: the oracle's BiCGSTAB(2)
        t0 = po.wall()
        cret, _ = po.bicgstabl(n, rp, ci, v, b, 2, 1e-10, 1000)
        cpu = {"value": round((po.wall() - t0) * 1e3, 3), "unit": "ms", "cores": 1, "kind": "port",
               "outer_iters": cret}
    return {"workload": "BiCGSTAB(2) 64x64 convection-diffusion, n=4096, tol 1e-10",
            "gpu_ms": round(ms, 3), "outer_iters": info.ret, "true_res": info.true_res,
            "path": "one persistent workgroup", "multi_kernel_graph_ms": round(ms_g, 3),
            "per_outer_iter_us": round(ms * 1e3 / max(1, info.ret + 1), 1), "cpu_baseline": cpu}


def bench_cfg2_block(dv, gen, m=128, s=8):
    """Block CG (s = 8) on the 128^3 Poisson matrix, tol 1e-8: plain FP64 and FP32 inner
    iterations with FP64 reliable updates, with the reference's first direction P0 = 2V
    and with P0 = V."""
    n, rp, ci, v = gen.poisson3d(m, lower=True)
    B = gen.normal(202, n * s).reshape(s, n)            # column-major m x s: column j = B[j]
    M = dv.Matrix(n, rp, ci, v, symmetric=True, fp32=True)
    dB, dX = dv.DeviceArray.from_numpy(B), dv.DeviceArray((s, n))

    def run(fix, prec=0):
        dX.zero()
        return dv.blcg(M, s, dB, dX, 1e-8, 5000, fix_blcg_p0=fix, precision=prec)

    run(0)
    out = {"workload": f"block CG s={s}, Poisson {m}^3, tol 1e-8"}
    for fix, key in ((0, "p0_2v_as_reference"), (1, "p0_v")):
        info, ms = _timed(lambda: run(fix), 1)
        X64 = dX.numpy()
        run(fix, dv.BKH_FP32_IR)
        i32, ms32 = _timed(lambda: run(fix, dv.BKH_FP32_IR), 1)
        out[key] = {"fp64_ms": round(ms, 2), "iters": info.ret, "true_res": info.true_res,
                    "ms_per_iter": round(ms / max(1, info.ret + 1), 3),
                    "fp32_ir": {"ms": round(ms32, 2), "inner_iters": i32.inner_iters, "passes": i32.passes,
                                "true_res": i32.true_res, "speedup_vs_fp64": round(ms / ms32, 3),
                                "x_rel_err_vs_fp64": _rel(dX.numpy(), X64)}}
    return out


def bench_cfg3(dv, gen, m=96):
    """Shifted BiCGSTAB(4), 96^3 convection-diffusion, 64 shifts, tol 1e-8 (seed)."""
    px, py, pz = gen.cfg3_peclet()
    n, rp, ci, v = gen.convdiff3d(m, px, py, pz)
    b = gen.normal(303, n)
    sig = 10.0 ** (-3 + 3 * np.arange(64) / 63.0)
    M = dv.Matrix(n, rp, ci, v)
    db, dx = dv.DeviceArray.from_numpy(b), dv.DeviceArray((n * 65,))
    def run(**kw):
        dx.zero()
        return dv.shbicgstabl(M, db, 4, sig, dx, 1e-8, 1000, **kw)

    run()
    (info, tr), ms = _timed(run, 1)
    X64 = dx.numpy().reshape(65, n)
    (info32, tr32), ms32 = _timed(lambda: run(precision=dv.BKH_FP32_IR), 1)
    X32 = dx.numpy().reshape(65, n)
    err = max(_rel(X32[q], X64[q]) for q in range(65))
    return {"workload": f"shifted BiCGSTAB(4), convection-diffusion {m}^3, n={n}, 64 shifts, tol 1e-8",
            "fp64_ms": round(ms, 2), "outer_iters": info.ret, "max_true_res": float(tr.max()),
            "fp32_ir": {"ms": round(ms32, 2), "outer_iters": info32.ret, "max_true_res": float(tr32.max()),
                        "speedup_vs_fp64": round(ms / ms32, 3), "max_x_rel_err_vs_fp64": err}}


def bench_complex(dv, gen, po, m=96, ns=16):
    """Complex FP64 family: shifted BiCGSTAB(4), BiCGSTAB(4) and block BiCGSTAB(4) (s = 4,
    RHS seeds 305/306) on the 96^3
    convection-diffusion operator with an imaginary diagonal shift (i 0.05 a_ii),
    16 complex shifts sigma_k = 10^(-2 + 2k/15) (1 + 0.5i), RHS N(0,1)+iN(0,1) seeds
    303/304, tol 1e-8; complex SpMV roofline; oracle CPU baseline on a bounded sample."""
    n, rp, ci, v = gen.zconvdiff3d(m)
    nnz = int(rp[-1])
    b = gen.normal(303, n) + 1j * gen.normal(304, n)
    sig = 10.0 ** (-2 + 2 * np.arange(ns) / (ns - 1.0)) * (1 + 0.5j)
    M = dv.ZMatrix(n, rp, ci, v)
    db = dv.DeviceArray.from_numpy(b)
    dxs = dv.DeviceArray((n * (ns + 1),), np.complex128)
    dx = dv.DeviceArray((n,), np.complex128)

    def shifted():
        dxs.zero()
        return dv.zshbicgstabl(M, db, 4, sig, dxs, 1e-8, 1000)

    def single():
        dx.zero()
        return dv.zbicgstabl(M, db, 4, dx, 1e-8, 1000)

    shifted()
    (sinfo, tr), sms = _timed(shifted, 1)
    single()
    info, ms = _timed(single, 1)
    sb = 4
    B = (gen.normal(305, n * sb) + 1j * gen.normal(306, n * sb)).reshape(sb, n)     # column-major n x sb
    dB = dv.DeviceArray.from_numpy(B.ravel())
    dX = dv.DeviceArray((n * sb,), np.complex128)

    def block():
        dX.zero()
        return dv.zblbicgstabl(M, sb, dB, 4, dX, 1e-8, 1000)

    block()
    binfo, bms = _timed(block, 1)
    M.bench_spmv(reps=3)
    spmv_ms = M.bench_spmv(reps=20)
    alg = nnz * (16 + 4) + 4 * (n + 1) + 16 * n + 16 * n
    gbs = alg / (spmv_ms * 1e-3) / 1e9
    out = {"workload": f"complex shifted BiCGSTAB(4), BiCGSTAB(4) and block BiCGSTAB(4) (s={sb}), convection-diffusion "
                       f"{m}^3 + i 0.05 diag, n={n}, nnz={nnz}, {ns} complex shifts, tol 1e-8",
           "zshbicgstabl_ms": round(sms, 2), "zshbicgstabl_outer_iters": sinfo.ret,
           "zshbicgstabl_max_true_res": float(tr.max()),
           "zbicgstabl_ms": round(ms, 2), "zbicgstabl_outer_iters": info.ret, "zbicgstabl_true_res": info.true_res,
           "zblbicgstabl_s4_ms": round(bms, 2), "zblbicgstabl_outer_iters": binfo.ret,
           "zblbicgstabl_true_res": binfo.true_res,
           "roofline": {"bound": "hbm", "kernel": "k_zspmv<false> (complex SELL-64, lane per row)",
                        "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(gbs / HBM_PEAK_GBS, 4), "algorithmic_bytes": alg,
                        "avg_launch_us": round(spmv_ms * 1e3, 2)}}
    if po is not None:
        # bounded sample: 2 outer iterations of the oracle's restatements, scaled to
        # the device's outer iteration counts
        rpi, cii = np.ascontiguousarray(rp, np.int32), np.ascontiguousarray(ci, np.int32)
        t0 = time.perf_counter()
        po.zbicgstabl(n, rpi, cii, v, b, 4, 1e-300, 2)
        per1 = (time.perf_counter() - t0) / 2
        t0 = time.perf_counter()
        po.zshbicgstabl(n, rpi, cii, v, b, 4, sig, 1e-300, 1)
        pers = time.perf_counter() - t0
        t0 = time.perf_counter()
        po.zblbicgstabl(n, rpi, cii, v, B.T, 4, 1e-300, 1)
        perb = time.perf_counter() - t0
        out["cpu_baseline"] = {
            "zbicgstabl_ms": round(per1 * (info.ret + 1) * 1e3, 1), "zshbicgstabl_ms": round(pers * (sinfo.ret + 1) * 1e3, 1),
            "zblbicgstabl_s4_ms": round(perb * (binfo.ret + 1) * 1e3, 1),
            "unit": "ms", "cores": 1, "kind": "port",
            "sample": f"oracle (C restatement of src/bicgstabl.c:141-276, src/shbicgstabl.c:288-592 and "
                      f"src/blbicgstabl.c:255-486, gcc -O3, "
                      f"1 thread): 2 outer iterations of zbicgstabl ({per1 * 1e3:.0f} ms each), 1 of "
                      f"zshbicgstabl ({pers * 1e3:.0f} ms) and 1 of zblbicgstabl s=4 ({perb * 1e3:.0f} ms), "
                      f"scaled to the device's outer iteration counts"}
    return out


def bench_cfg4(dv, gen):
    """BiCGSTAB(2) on 512 independent uneven systems, tol 1e-8: batched (one launch per
    phase, autotuned layout) FP32-IR and FP64 against 512 sequential plain FP64 solves."""
    systems = gen.cfg4_batch()
    sizes = [sy[0] for sy in systems]
    Bt = dv.Batch(systems, fp32=True)
    tune_ms = Bt.tune(fp32=True)
    b = gen.normal(405, Bt.N)
    db, dx = dv.DeviceArray.from_numpy(b), dv.DeviceArray((Bt.N,))
    def run(**kw):
        dx.zero()
        return Bt.bicgstabl(db, 2, dx, 1e-8, 1000, **kw)

    run(precision=dv.BKH_FP32_IR)
    (i32, r32, t32), ms32 = _timed(lambda: run(precision=dv.BKH_FP32_IR), 2)
    x32 = dx.numpy()
    (i64, r64, t64), ms64 = _timed(lambda: run(), 2)
    x64 = dx.numpy()
    off0 = Bt.offsets
    err32 = max(_rel(x32[off0[q]:off0[q + 1]], x64[off0[q]:off0[q + 1]]) for q in range(len(sizes)))
    # the plain FP64 path: one bkh_bicgstabl per system, in sequence
    mats, bs, xs = [], [], []
    off = Bt.offsets
    for q, sy in enumerate(systems):
        mats.append(dv.Matrix(*sy))
        bs.append(dv.DeviceArray.from_numpy(b[off[q]:off[q + 1]]))
        xs.append(dv.DeviceArray((sy[0],)))
    del systems

    def seq():
        worst = 0.0
        for q in range(len(mats)):
            xs[q].zero()
            inf = dv.bicgstabl(mats[q], bs[q], 2, xs[q], 1e-8, 1000)
            worst = max(worst, inf.true_res)
        return worst

    worst_seq, ms_seq = _timed(seq, 1)
    # the whole solve as one persistent cooperative launch (bitwise the graph path's x);
    # one untimed call first, as for the graph path (first launch of the kernel)
    run(precision=dv.BKH_FP32_IR, path=2)
    run(path=2)
    (ip32, rp32, tp32), msp32 = _timed(lambda: run(precision=dv.BKH_FP32_IR, path=2), 1)
    same32 = bool(np.array_equal(dx.numpy(), x32))
    (ip64, rp64, tp64), msp64 = _timed(lambda: run(path=2), 1)
    same64 = bool(np.array_equal(dx.numpy(), x64))
    layout = Bt.layout()
    return {"workload": f"512 uneven systems n in [2^10, 2^17], N={Bt.N}, nnz={Bt.nnz}, BiCGSTAB(2), tol 1e-8",
            "sequential_fp64_ms": round(ms_seq, 1), "sequential_max_true_res": worst_seq,
            "batched_fp64_ms": round(ms64, 1), "batched_fp64_max_true_res": float(t64.max()),
            "batched_fp32_ir_ms": round(ms32, 1), "batched_fp32_ir_max_true_res": float(t32.max()),
            "batched_fp32_ir_max_x_rel_err_vs_fp64": err32,
            "fp32_ir_passes": i32.passes, "converged": int((r32 >= 0).sum()),
            "speedup_vs_sequential_fp64": round(ms_seq / ms32, 2),
            "speedup_vs_batched_fp64": round(ms64 / ms32, 3),
            "persistent_launch": {"fp32_ir_ms": round(msp32, 1), "fp64_ms": round(msp64, 1),
                                  "x_equals_graph_path": same32 and same64,
                                  "max_true_res": max(float(tp32.max()), float(tp64.max()))},
            "autotuned_layout": {"spmv_kernel": "column-chunked" if Bt.spmv_kernel else "csr-stream",
                                 "rows_per_wave": layout[0], "nnz_per_wave": layout[1],
                                 "tile_rows": layout[2], "fp32_spmv_ms": round(tune_ms, 3)}}


def bench_tuner(dv, gen):
    """bkh_matrix_tune on the 128^3 Poisson matrix and on a 2M-row random-column matrix:
    the SpMV variant / wave blocking and bkh_spmm columns per pass it keeps, against the
    creation-time choice; and bkh_spmm with s = 32 columns (chunks of the tuned width)."""
    out = {}
    for name, (n, rp, ci, v), sym in (("poisson128", gen.poisson3d(128, lower=True), True),
                                      ("random_rows_2M", gen.random_rows(6060, 1 << 21), False)):
        M = dv.Matrix(n, rp, ci, v, symmetric=sym)
        r = M.tune(16)
        names = {0: "workgroup CSR-stream", 1: "wave CSR-stream", 2: "SELL-64", 3: "SELL-64 int32 columns"}
        X = dv.DeviceArray.from_numpy(gen.normal(9, n * 32))
        Y = dv.DeviceArray((n * 32,))
        M.spmm(32, X, Y)
        _, ms32 = _timed(lambda: M.spmm(32, X, Y), 3)
        out[name] = {"n": n, "nnz": M.nnz, "spmv": names[r["spmv_variant"]], "rows_per_wave": r["rows_per_wave"],
                     "nnz_per_wave": r["nnz_per_wave"], "candidates": r["candidates"],
                     "spmv_ms": round(r["spmv_ms"], 4), "spmv_ms_creation_choice": round(r["spmv_ms_default"], 4),
                     "spmm_width": {s_: r["spmm_width"][s_] for s_ in (2, 4, 8, 12, 16)},
                     "spmm_ms": {s_: round(r["spmm_ms"][s_], 4) for s_ in (2, 4, 8, 12, 16)},
                     "spmm_ms_one_pass": {s_: round(r["spmm_ms_full"][s_], 4) for s_ in (2, 4, 8, 12, 16)},
                     "spmm_s32_ms_incl_host_sync": round(ms32, 3)}
        del M, X, Y
    return out


def bench_cfg5(dv, gen, n_train=200, m_train=64, n_test=50, m_test=128, n_plain=2, epochs=4):
    """Learned-FSAI PCG on variable-coefficient 3D diffusion (kappa = exp(GRF), sigma 2,
    correlation 8 cells): train the edge MLP on the device over n_train seeded fields
    (m_train^3 grids, same statistics), then solve the n_test held-out m_test^3 fields
    with Jacobi PCG, the untrained network's factor and the trained network's factor
    (and unpreconditioned CG on n_plain of them)."""
    from concurrent.futures import ThreadPoolExecutor
    from bksp_amd import fsai
    t0 = time.perf_counter()
    seeds = list(range(5000, 5000 + n_train))
    with ThreadPoolExecutor(16) as ex:                 # host GRF generation (C, GIL released)
        systems = list(ex.map(lambda s: fsai.field_system(s, m_train), seeds))
    fields = [dv.FsaiField(n, *low, Z=fsai.probes(n, low, 10 * s + 1)) for s, (n, low) in zip(seeds, systems)]
    del systems
    prep_s = time.perf_counter() - t0
    model = dv.FsaiModel(fsai.init_params())
    t0 = time.perf_counter()
    hist = fsai.train(model, fields, epochs=epochs, lr=3e-3, lr_final=3e-4)
    from bksp_amd import load
    load().bkh_sync()
    train_s = time.perf_counter() - t0
    for f in fields:
        f.destroy()
    untrained = dv.FsaiModel(fsai.init_params())
    rows = []
    with ThreadPoolExecutor(8) as ex:                  # held-out fields generated ahead on host threads
        futs = [ex.submit(fsai.field_system, 6000 + q, m_test) for q in range(n_test)]
        for q in range(n_test):
            system = futs[q].result()
            futs[q] = None
            rows.append(fsai.solve_compare(model, 6000 + q, m_test, plain=q < n_plain, system=system,
                                           untrained=untrained))
            del system
    ui = np.array([r["untrained"]["iters"] for r in rows]); ums = np.array([r["untrained"]["ms"] for r in rows])
    jac = np.array([r["jacobi"]["ms"] for r in rows]); lea = np.array([r["learned"]["ms"] for r in rows])
    ji = np.array([r["jacobi"]["iters"] for r in rows]); li = np.array([r["learned"]["iters"] for r in rows])
    plain = [r["plain_cg"] for r in rows if "plain_cg" in r]
    pl_ms = float(np.mean([p["ms"] for p in plain])) if plain else None
    lea_p = float(np.mean(lea[:len(plain)])) if plain else None
    # the MLP forward on the held-out 128^3 fields: the median over fields of each
    # field's factor build (the forward, ~95 %, and the scatter of G and G^T to their
    # slice-diagonal and SELL copies), 3 timed builds per field
    edges = rows[0]["learned"]["edges"]
    fwd_ms = float(np.median([r["learned"]["factor_build_ms"] for r in rows]))
    flops = 2 * (8 * 64 + 64 * 64 + 64) * edges
    return {"workload": f"learned-FSAI PCG, FV diffusion kappa=exp(GRF sigma 2, corr 8), "
                        f"train {n_train} fields {m_train}^3, test {n_test} held-out fields {m_test}^3, tol 1e-8",
            "train": {"prep_s": round(prep_s, 1), "train_s": round(train_s, 2), "steps": epochs * n_train,
                      "loss_first_epoch": round(hist[0], 5), "loss_last_epoch": round(hist[-1], 5)},
            "mlp_forward": {"edges": edges, "field": f"{m_test}^3", "ms": round(fwd_ms, 4),
                            "TFLOPs": round(flops / (fwd_ms * 1e-3) / 1e12, 2),
                            "bound": "valu (the activations: 627 VALU against 12 MFMA instructions per "
                                     "16-edge tile, profiles/fsai_mlp_field128.md)",
                            "frac_of_2500": round(flops / (fwd_ms * 1e-3) / 1e12 / 2500.0, 4)},
            "jacobi": {"mean_iters": float(ji.mean()), "mean_ms": round(float(jac.mean()), 2),
                       "max_true_res": max(r["jacobi"]["true_res"] for r in rows)},
            "learned": {"mean_iters": float(li.mean()), "mean_ms": round(float(lea.mean()), 2),
                        "max_true_res": max(r["learned"]["true_res"] for r in rows),
                        "mean_factor_build_ms": round(float(np.mean([r["learned"]["factor_build_ms"] for r in rows])), 3)},
            "plain_cg": {"fields": len(plain), "mean_ms": round(pl_ms, 1) if pl_ms else None,
                         "mean_iters": float(np.mean([p["iters"] for p in plain])) if plain else None,
                         "max_true_res": max(p["true_res"] for p in plain) if plain else None},
            "untrained": {"mean_iters": float(ui.mean()), "mean_ms": round(float(ums.mean()), 2),
                          "max_true_res": max(r["untrained"]["true_res"] for r in rows)},
            "learned_iters_vs_untrained": round(float(li.mean() / ui.mean()), 3),
            "speedup_learned_vs_untrained": round(float(ums.mean() / lea.mean()), 3),
            "learned_iters_vs_jacobi": round(float(li.mean() / ji.mean()), 3),
            "speedup_learned_vs_jacobi": round(float(jac.mean() / lea.mean()), 3),
            "speedup_learned_vs_plain_fp64_cg": round(pl_ms / lea_p, 2) if plain else None}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--m", type=int, default=128)
    ap.add_argument("--configs", default="cfg1,cfg2b,cfg3,cfg4,cfg5,complex,tuner",
                    help="extra workloads reported beside the main line ('' for none)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU "
              "row partition is a design note, not part of this build.", file=sys.stderr)
        sys.exit(2)

    from bksp_amd import device as dv
    from bksp_amd import gen

    m = args.m
    n, rp, ci, v = gen.poisson3d(m, lower=True)
    b = gen.normal(202, n)
    M = dv.Matrix(n, rp, ci, v, symmetric=True, fp32=True)
    db = dv.DeviceArray.from_numpy(b)
    dx = dv.DeviceArray((n,))
    tol = 1e-8

    def solve(**kw):
        dx.zero()                       # x0 = 0 on the device (no host transfer in the timed loop)
        return dv.cg(M, db, dx, tol, 20000, **kw)

    for _ in range(args.warmup):
        solve()
    from bksp_amd import load
    load().bkh_sync()
    t0 = time.perf_counter()
    infos, each = [], []
    for _ in range(args.steps):          # each solve returns after its device work ends
        t1 = time.perf_counter()
        infos.append(solve())
        each.append((time.perf_counter() - t1) * 1e3)
    load().bkh_sync()
    wall_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    info = infos[-1]
    dev_ms = float(np.mean([i.time_ms for i in infos]))
    x64 = dx.numpy()

    # FP32-inner / FP64-refinement path on the same system (secondary line): the
    # median of at least 5 solves (its reliable updates wait on the host 4-5 times per
    # solve, so a single slow host round trip moves a mean), against the FP64 median
    solve(precision=dv.BKH_FP32_IR)
    ir, ir_each = [], []
    for _ in range(max(5, args.steps)):
        t1 = time.perf_counter()
        ir.append(solve(precision=dv.BKH_FP32_IR))
        ir_each.append((time.perf_counter() - t1) * 1e3)
    ir_wall = float(np.median(ir_each))
    fp64_median = float(np.median(each))
    x32 = dx.numpy()

    # the bksp_dcg drop-in on host arrays: the first call uploads A and builds its device
    # layout (expansion of the stored triangle, SELL-64 and slice-diagonal copies); later
    # calls with the same dspmat hit the per-thread device-copy cache (contents hashed)
    # and pay the hash, the b / x transfers and the solve
    from bksp_amd import bksp
    Ah = bksp.dspmat(n, rp, ci, v, symmetric=True)
    xh = np.zeros(n)
    t1 = time.perf_counter()
    dret = bksp.bksp_dcg(Ah, b, tol, 20000, xh)
    drop_first = (time.perf_counter() - t1) * 1e3
    drop_each = []
    for _ in range(3):
        xh[:] = 0.0
        t1 = time.perf_counter()
        bksp.bksp_dcg(Ah, b, tol, 20000, xh)
        drop_each.append((time.perf_counter() - t1) * 1e3)
    dropin = {"first_call_ms": round(drop_first, 2), "cached_call_ms": round(float(np.median(drop_each)), 2),
              "ret": dret, "x_equals_device_path": bool(np.array_equal(xh, x64)),
              "note": "host arrays in and out (PCIe-inclusive); not value"}
    del Ah, xh

    # dominant kernel of the timed solve: CG's fused kernel A (SELL SpMV with the
    # direction update p = r + beta p in its gather), timed with HIP events on its
    # stream; the plain SpMV beside it.
    nnz = int(M.nnz)
    M.bench_cg_kernel(reps=5)
    cga_ms = M.bench_cg_kernel(reps=50)
    cga_alg = 12 * nnz + 4 * (n + 1) + 32 * n           # A, row pointers, p_old and r read, q and p_new written
    if M.dia_entries > 0:   # as stored: slice-diagonal values, per-slice offsets and slice offsets
        cga_fmt = 8 * M.dia_entries + 4 * (M.dia_entries // 64) + 4 * ((n + 63) // 64 + 1) + 32 * n
        cga_kernel = ("k_cgf_a_sell<double,double,int,1024,DIA> (CG: SpMV on the slice-diagonal copy with "
                      "p = r + beta p in the gather)")
    else:                   # as stored (SELL entries, 16-bit column deltas)
        cga_fmt = M.spmv_bytes() - 16 * n + 32 * n
        cga_kernel = "k_cgf_a_sell<double,double,short,1024> (CG: SELL-64 SpMV with p = r + beta p in the gather)"
    cga_gbs = cga_alg / (cga_ms * 1e-3) / 1e9
    # the same kernel on a matrix whose working set exceeds the 256 MB Infinity Cache
    # (192^3 Poisson: 395 MB of values alone), so repeated launches stream from HBM
    nb_, rpb, cib, vb = gen.poisson3d(192, lower=True)
    Mb = dv.Matrix(nb_, rpb, cib, vb, symmetric=True)
    del rpb, cib, vb
    Mb.bench_cg_kernel(reps=5)
    big_ms = Mb.bench_cg_kernel(reps=50)
    big_alg = 12 * int(Mb.nnz) + 4 * (nb_ + 1) + 32 * nb_
    big_fmt = (8 * Mb.dia_entries + 4 * (Mb.dia_entries // 64) + 4 * ((nb_ + 63) // 64 + 1) + 32 * nb_
               if Mb.dia_entries > 0 else None)
    beyond = {"workload": f"Poisson 192^3, n={nb_}, nnz={int(Mb.nnz)} (working set > 256 MB Infinity Cache)",
              "avg_launch_us": round(big_ms * 1e3, 2), "algorithmic_bytes": big_alg,
              "achieved": round(big_alg / (big_ms * 1e-3) / 1e9, 1),
              "frac": round(big_alg / (big_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
              "format_bytes": big_fmt,
              "format_frac": round(big_fmt / (big_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) if big_fmt else None}
    Mb.destroy()
    del Mb
    dy = dv.DeviceArray((n,))
    dv.bench_spmv(M, db, dy, reps=5)
    spmv_ms = dv.bench_spmv(M, db, dy, reps=50)
    bytes64 = spmv_bytes(n, nnz)
    gbs = bytes64 / (spmv_ms * 1e-3) / 1e9
    fmt64 = M.spmv_bytes()
    db32 = dv.DeviceArray.from_numpy(b.astype(np.float32))
    dy32 = dv.DeviceArray((n,), np.float32)
    spmv32_ms = dv.bench_spmv(M, db32, dy32, reps=50, fp32=True)
    gbs32 = spmv_bytes(n, nnz, 4, 4) / (spmv32_ms * 1e-3) / 1e9
    big = 1 << 29
    ca, cb = dv.DeviceArray((big // 8,)), dv.DeviceArray((big // 8,))
    dv.bench_copy(ca, cb, big, reps=3)
    copy_gbs = 2 * big / (dv.bench_copy(ca, cb, big, reps=20) * 1e-3) / 1e9
    ca.free(); cb.free()

    t_cga = load_traffic("cg_fused_poisson128_traffic.json")
    t_spmv = load_traffic("spmv_fp64_poisson128_traffic.json")
    roofline = {"bound": "hbm", "kernel": cga_kernel,
                "achieved": round(cga_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(cga_gbs / HBM_PEAK_GBS, 4),
                "traffic": t_cga.get("bytes_per_launch") if t_cga else None,
                "algorithmic_bytes": cga_alg, "avg_launch_us": round(cga_ms * 1e3, 2),
                "format_bytes": cga_fmt, "format_GBs": round(cga_fmt / (cga_ms * 1e-3) / 1e9, 1),
                "format_frac": round(cga_fmt / (cga_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                "note": ("algorithmic bytes count a 4-byte column index per nonzero and the row pointers; "
                         "the slice-diagonal copy stores neither (per-slice column offsets), so frac on "
                         "algorithmic bytes can reach 1; format_frac is the rate on the bytes the kernel "
                         "stores, traffic the HBM bytes rocprof counts per launch")
                if M.dia_entries > 0 else None,
                "measured_copy_GBs": round(copy_gbs, 1),
                "frac_of_measured_copy": round(cga_gbs / copy_gbs, 4),
                "beyond_cache": beyond,
                "spmv": {"kernel": "k_spmv_sell<double,double,short> (SELL-64, 16-bit column deltas)",
                         "achieved": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4),
                         "algorithmic_bytes": bytes64, "avg_launch_us": round(spmv_ms * 1e3, 2),
                         "format_bytes": fmt64, "format_GBs": round(fmt64 / (spmv_ms * 1e-3) / 1e9, 1),
                         "format_frac": round(fmt64 / (spmv_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                         "traffic": t_spmv.get("bytes_per_launch") if t_spmv else None},
                "fp32_spmv": {"achieved": round(gbs32, 1), "avg_launch_us": round(spmv32_ms * 1e3, 2),
                              "algorithmic_bytes": spmv_bytes(n, nnz, 4, 4),
                              "format_bytes": M.spmv_bytes(fp32=True)}}

    cpu = None
    if not args.no_cpu:
        from oracle import pyoracle as po
        model, threads, avail = _host_info()
        t0 = po.wall()
        cret, cx = po.cg(n, rp, ci, v, b, tol, 20000)
        cpu_s = po.wall() - t0
        cpu = {"value": round(cpu_s * 1e3, 1), "unit": "ms", "cores": 1, "kind": "port",
               "sample": f"one complete solve of the same system by the oracle's CG (C restatement of "
                         f"src/cg.c, gcc -O3, one thread), {cret} iterations to 1e-8",
               "cpu_model": model, "host_cores_available": avail,
               "x_rel_diff_vs_gpu": _rel(x64, cx), "iterations": cret}
        if po.lib_mt() is not None:
            os.environ.setdefault("OMP_NUM_THREADS", str(threads))
            t0 = po.wall()
            mret, _ = po.cg_mt(n, rp, ci, v, b, tol, 20000)
            cpu["threaded"] = {"value": round((po.wall() - t0) * 1e3, 1), "unit": "ms", "cores": threads,
                               "iterations": mret,
                               "sample": "the same solve with the BLAS-1 restatements threaded (OpenMP) and "
                                         "the SpMV serial, as the reference runs with a multithreaded BLAS"}

    extra = {}
    want = [c for c in args.configs.split(",") if c]
    if want:
        po = None
        if not args.no_cpu:
            from oracle import pyoracle as po    # cpu_baseline legs only
        del M, db, dx, dy, db32, dy32
        for name, fn in (("cfg1", lambda: bench_cfg1(dv, gen, po)), ("cfg2_block", lambda: bench_cfg2_block(dv, gen)),
                         ("cfg3", lambda: bench_cfg3(dv, gen)), ("cfg4", lambda: bench_cfg4(dv, gen)),
                         ("cfg5", lambda: bench_cfg5(dv, gen)), ("complex", lambda: bench_complex(dv, gen, po)),
                         ("tuner", lambda: bench_tuner(dv, gen))):
            key = {"cfg2_block": "cfg2b"}.get(name, name)
            if key in want:
                try:
                    extra[name] = fn()
                except Exception as e:          # reported, never silently dropped
                    extra[name] = {"error": repr(e)}
                print(f"[bench] {name}: {json.dumps(extra[name])}", file=sys.stderr, flush=True)

    out = {
        "metric": METRIC, "value": round(wall_ms, 3), "unit": "ms", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(wall_ms, 3),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"CG to 1e-8 on 3D 7-point Poisson {m}^3 (bkh_cg on HBM-resident inputs; "
                               f"bksp_dcg drop-in in 'dropin'), n={n}, "
                               f"nnz={nnz} expanded ({int(rp[-1])} stored lower), RHS N(0,1) seed 202, x0=0",
                   "n": n, "nnz": nnz, "tol": tol,
                   "inputs_sha256": _sha256(rp, ci, v, b)},
        "iterations": info.ret, "true_res": info.true_res, "device_ms": round(dev_ms, 3),
        "fp64_median_ms": round(fp64_median, 3),
        "fp32_ir": {"ms": round(ir_wall, 3), "mean_ms": round(float(np.mean(ir_each)), 3), "solves": len(ir_each),
                    "inner_iters": ir[-1].inner_iters, "passes": ir[-1].passes,
                    "true_res": ir[-1].true_res, "speedup_vs_fp64": round(fp64_median / ir_wall, 3),
                    "x_rel_err_vs_fp64": _rel(x32, x64)},
        "dropin": dropin,
        "roofline": roofline, "cpu_baseline": cpu,
        "configs": extra,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
