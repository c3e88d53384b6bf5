"""Runs one complete solve of a benchmark workload (after one warm-up solve) so that
rocprofv3 kernel traces give the per-kernel time of a whole solve:

  BKH_EAGER=1 rocprofv3 --kernel-trace --stats -d OUT -- \
      python3 -m bksp_amd.profile_solvers --solver cg --precision fp32ir

BKH_EAGER=1 issues the iteration kernels as plain launches (the graph replays issue
the same kernels in the same order); the per-kernel durations are what the solve's
graphs run.

Workloads: cg / blcg (128^3 Poisson, RHS N(0,1) seed 202, s = 8 for block CG),
shifted (96^3 convection-diffusion, 64 shifts, l = 4), pcg (128^3 exp(GRF) diffusion
field 6000, Jacobi or an untrained learned factor), bicgstabl (64x64 convection-
diffusion, l = 2, tol 1e-10), zbicgstabl / zshbicgstabl (the complex 96^3
convection-diffusion + i 0.05 diag operator, l = 4, 16 complex shifts).
"""
import argparse
import json
import time

import numpy as np

from . import device as dv
from . import gen


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--solver", choices=["cg", "blcg", "shifted", "pcg", "bicgstabl", "zbicgstabl", "zshbicgstabl",
                                         "zblbicgstabl", "blbicgstabl", "batched", "batch_spmv"],
                    default="cg")
    ap.add_argument("--precision", choices=["fp64", "fp32ir"], default="fp64")
    ap.add_argument("--precond", choices=["jacobi", "fsai"], default="jacobi")
    ap.add_argument("--m", type=int, default=0)
    ap.add_argument("--path", type=int, default=0, help="batched: 2 = the whole solve as one persistent launch")
    args = ap.parse_args()
    prec = dv.BKH_FP32_IR if args.precision == "fp32ir" else dv.BKH_FP64
    if args.solver == "cg":
        m = args.m or 128
        n, rp, ci, v = gen.poisson3d(m, lower=True)
        M = dv.Matrix(n, rp, ci, v, symmetric=True, fp32=True)
        db, dx = dv.DeviceArray.from_numpy(gen.normal(202, n)), dv.DeviceArray((n,))

        def run():
            dx.zero()
            return dv.cg(M, db, dx, 1e-8, 20000, precision=prec)
    elif args.solver == "blcg":
        m = args.m or 128
        n, rp, ci, v = gen.poisson3d(m, lower=True)
        M = dv.Matrix(n, rp, ci, v, symmetric=True, fp32=True)
        dB, dX = dv.DeviceArray.from_numpy(gen.normal(202, n * 8).reshape(8, n)), dv.DeviceArray((8, n))

        def run():
            dX.zero()
            return dv.blcg(M, 8, dB, dX, 1e-8, 5000, precision=prec)
    elif args.solver == "shifted":
        m = args.m or 96
        px, py, pz = gen.cfg3_peclet()
        n, rp, ci, v = gen.convdiff3d(m, px, py, pz)
        M = dv.Matrix(n, rp, ci, v, fp32=True)
        sig = 10.0 ** (-3 + 3 * np.arange(64) / 63.0)
        db, dx = dv.DeviceArray.from_numpy(gen.normal(303, n)), dv.DeviceArray((n * 65,))

        def run():
            dx.zero()
            return dv.shbicgstabl(M, db, 4, sig, dx, 1e-8, 1000, precision=prec)[0]
    elif args.solver == "pcg":
        from . import fsai
        m = args.m or 128
        n, (rpl, cil, vl) = fsai.field_system(6000, m)
        F = dv.FsaiField(n, rpl, cil, vl)
        db, dx = dv.DeviceArray.from_numpy(gen.normal(505, n)), dv.DeviceArray((n,))
        if args.precond == "fsai":
            F.factor(dv.FsaiModel(fsai.init_params(3) * 2))

            def run():
                dx.zero()
                return F.pcg(db, dx, 1e-8, 100000)
        else:
            def run():
                dx.zero()
                return F.pcg_plain(1, db, dx, 1e-8, 100000)
    elif args.solver in ("batched", "batch_spmv"):
        systems = gen.cfg4_batch()
        Bt = dv.Batch(systems, fp32=True)
        del systems
        Bt.tune(fp32=prec == dv.BKH_FP32_IR)
        n = Bt.N
        db, dx = dv.DeviceArray.from_numpy(gen.normal(405, n)), dv.DeviceArray((n,))
        if args.solver == "batch_spmv":            # 20 SpMVs in the solve's precision
            dt = np.float32 if prec == dv.BKH_FP32_IR else np.float64
            xs, ys = dv.DeviceArray.from_numpy(gen.normal(406, n).astype(dt)), dv.DeviceArray((n,), dt)

            class _Info:
                ret = inner_iters = passes = 0
                time_ms = true_res = 0.0

            def run():
                for _ in range(20):
                    Bt.spmv(xs, ys, fp32=prec == dv.BKH_FP32_IR)
                return _Info()
        else:
            def run():
                dx.zero()
                return Bt.bicgstabl(db, 2, dx, 1e-8, 1000, precision=prec, path=args.path)[0]
    elif args.solver == "blbicgstabl":
        m = args.m or 96
        px, py, pz = gen.cfg3_peclet()
        n, rp, ci, v = gen.convdiff3d(m, px, py, pz)
        M = dv.Matrix(n, rp, ci, v)
        sb = 4
        dB = dv.DeviceArray.from_numpy(gen.normal(307, n * sb))
        dx = dv.DeviceArray((n * sb,))

        def run():
            dx.zero()
            return dv.blbicgstabl(M, sb, dB, 4, dx, 1e-8, 1000)
    elif args.solver == "zblbicgstabl":
        m = args.m or 96
        n, rp, ci, v = gen.zconvdiff3d(m)
        M = dv.ZMatrix(n, rp, ci, v)
        sb = 4
        B = (gen.normal(305, n * sb) + 1j * gen.normal(306, n * sb)).reshape(sb, n)
        dB = dv.DeviceArray.from_numpy(B.ravel())
        dx = dv.DeviceArray((n * sb,), np.complex128)

        def run():
            dx.zero()
            return dv.zblbicgstabl(M, sb, dB, 4, dx, 1e-8, 1000)
    elif args.solver in ("zbicgstabl", "zshbicgstabl"):
        m = args.m or 96
        n, rp, ci, v = gen.zconvdiff3d(m)
        M = dv.ZMatrix(n, rp, ci, v)
        db = dv.DeviceArray.from_numpy(gen.normal(303, n) + 1j * gen.normal(304, n))
        sig = 10.0 ** (-2 + 2 * np.arange(16) / 15.0) * (1 + 0.5j)
        if args.solver == "zbicgstabl":
            dx = dv.DeviceArray((n,), np.complex128)

            def run():
                dx.zero()
                return dv.zbicgstabl(M, db, 4, dx, 1e-8, 1000)
        else:
            dx = dv.DeviceArray((n * 17,), np.complex128)

            def run():
                dx.zero()
                return dv.zshbicgstabl(M, db, 4, sig, dx, 1e-8, 1000)[0]
    else:
        m = args.m or 64
        n, rp, ci, v = gen.convdiff2d(m, 10.0, 20.0)
        import scipy.sparse as sp
        b = sp.csr_matrix((v, ci, rp), shape=(n, n)) @ gen.uniform(101, n, -1.0, 1.0)
        M = dv.Matrix(n, rp, ci, v)
        db, dx = dv.DeviceArray.from_numpy(b), dv.DeviceArray((n,))

        def run():
            dx.zero()
            return dv.bicgstabl(M, db, 2, dx, 1e-10, 1000)
    run()
    from . import load
    load().bkh_sync()
    t0 = time.perf_counter()
    info = run()
    load().bkh_sync()
    print(json.dumps({"solver": args.solver, "precision": args.precision, "n": int(n), "ret": info.ret,
                      "inner_iters": info.inner_iters, "passes": info.passes, "device_ms": round(info.time_ms, 3),
                      "wall_ms": round((time.perf_counter() - t0) * 1e3, 3), "true_res": info.true_res}), flush=True)


if __name__ == "__main__":
    main()
