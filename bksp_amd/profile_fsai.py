"""Runs the learned-FSAI MLP kernels on one 128^3 held-out field so rocprofv3 kernel
traces and counters can be read per dispatch: `--reps` factor builds (k_fsai_mlp
forward over every edge of tril(A)) and `--train` Adam steps (forward + backward).

  rocprofv3 --kernel-trace --stats -d OUT -- python3 -m bksp_amd.profile_fsai
  rocprofv3 --pmc SQ_INSTS_VALU SQ_INSTS_MFMA ... -d OUT -- python3 -m bksp_amd.profile_fsai
"""
import argparse

from . import device as dv
from . import fsai


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--m", type=int, default=128)
    ap.add_argument("--reps", type=int, default=10)
    ap.add_argument("--train", type=int, default=3)
    args = ap.parse_args()
    n, low = fsai.field_system(6000, args.m)
    F = dv.FsaiField(n, *low, Z=fsai.probes(n, low, 60001, residual_iters=()))
    M = dv.FsaiModel(fsai.init_params() * 2)
    ms = F.bench_forward(M, reps=args.reps)
    for _ in range(args.train):
        F.train_step(M, 1e-3)
    flops = 2 * (8 * 64 + 64 * 64 + 64) * F.nnzl
    print(f"field m={args.m} n={n} edges={F.nnzl}: forward avg_ms={ms:.4f} "
          f"({flops / (ms * 1e-3) / 1e12:.1f} TFLOP/s); {args.train} training steps", flush=True)


if __name__ == "__main__":
    main()
