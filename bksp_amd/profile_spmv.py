"""Runs one kernel on the 128^3 Poisson matrix `--reps` times -- the SpMV
(`--kernel spmv`, FP64 or `--fp32`) or CG's fused SpMV + direction-update kernel
(`--kernel cg`) -- plus a device copy of a known size, so that rocprofv3 kernel
traces and PMC counters (FETCH_SIZE, WRITE_SIZE) can be read per dispatch:

  rocprofv3 --kernel-trace --stats -d OUT -- python3 -m bksp_amd.profile_spmv
  rocprofv3 --pmc FETCH_SIZE -d OUT -- python3 -m bksp_amd.profile_spmv --kernel cg
"""
import argparse

import numpy as np

from . import device as dv
from . import gen


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--m", type=int, default=128)
    ap.add_argument("--reps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=0, help="untimed launches first (bench.py uses 5)")
    ap.add_argument("--fp32", action="store_true")
    ap.add_argument("--kernel", choices=["spmv", "cg"], default="spmv")
    args = ap.parse_args()
    n, rp, ci, v = gen.poisson3d(args.m, lower=True)
    M = dv.Matrix(n, rp, ci, v, symmetric=True, fp32=args.fp32)
    dt = np.float32 if args.fp32 else np.float64
    dx = dv.DeviceArray.from_numpy(gen.normal(202, n).astype(dt))
    dy = dv.DeviceArray((n,), dt)
    if args.kernel == "cg":
        if args.warmup:
            M.bench_cg_kernel(reps=args.warmup, fp32=args.fp32)
        ms = M.bench_cg_kernel(reps=args.reps, fp32=args.fp32)
    else:
        if args.warmup:
            dv.bench_spmv(M, dx, dy, reps=args.warmup, fp32=args.fp32)
        ms = dv.bench_spmv(M, dx, dy, reps=args.reps, fp32=args.fp32)
    nbytes = 1 << 28                                   # 256 MiB copy: 256 MiB read + 256 MiB written
    a, b = dv.DeviceArray((nbytes // 8,)), dv.DeviceArray((nbytes // 8,))
    cms = dv.bench_copy(a, b, nbytes, reps=5)
    print(f"{args.kernel} n={n} nnz={M.nnz} blocks={M.nblocks} avg_ms={ms:.4f}; copy {nbytes} B avg_ms={cms:.4f}",
          flush=True)


if __name__ == "__main__":
    main()
