"""Runs the FP64 CSR SpMV on the 128^3 Poisson matrix `--reps` times, plus a
device copy of a known size, so that rocprofv3 kernel traces and PMC counters
(FETCH_SIZE, WRITE_SIZE) can be read per dispatch:

  rocprofv3 --kernel-trace --stats -d OUT -- python -m bksp_amd.profile_spmv
  rocprofv3 --pmc FETCH_SIZE -d OUT -- python -m bksp_amd.profile_spmv
"""
import argparse

import numpy as np

from . import device as dv
from . import gen


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--m", type=int, default=128)
    ap.add_argument("--reps", type=int, default=20)
    ap.add_argument("--fp32", action="store_true")
    args = ap.parse_args()
    n, rp, ci, v = gen.poisson3d(args.m, lower=True)
    M = dv.Matrix(n, rp, ci, v, symmetric=True, fp32=args.fp32)
    dt = np.float32 if args.fp32 else np.float64
    dx = dv.DeviceArray.from_numpy(gen.normal(202, n).astype(dt))
    dy = dv.DeviceArray((n,), dt)
    ms = dv.bench_spmv(M, dx, dy, reps=args.reps, fp32=args.fp32)
    nbytes = 1 << 28                                   # 256 MiB copy: 256 MiB read + 256 MiB written
    a, b = dv.DeviceArray((nbytes // 8,)), dv.DeviceArray((nbytes // 8,))
    cms = dv.bench_copy(a, b, nbytes, reps=5)
    print(f"spmv n={n} nnz={M.nnz} blocks={M.nblocks} avg_ms={ms:.4f}; copy {nbytes} B avg_ms={cms:.4f}",
          flush=True)


if __name__ == "__main__":
    main()
