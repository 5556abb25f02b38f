"""Micro-benchmark of the SELL-64 block SpMV (drape_bsr_spmv) on an assembled cloth matrix.

Builds the 256x256 sheet (65,536 particles, 845,836 blocks), assembles its Hessian at rest on the GPU, and runs
y = A x `--reps` times back to back on the solver's SpMV kernel; prints the event-timed mean per launch and the
algorithmic bytes (40 B per block for values and column index, 4 B per row of row data, 12 B per row for each of
x and y) as one JSON line. Run it under rocprofv3 --kernel-trace for the kernel's own duration.

    python -m drape_amd.spmv_bench --reps 200
"""
from __future__ import annotations

import argparse
import json
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

from drape_amd import HipSolver, bsr_spmv, scenes  # noqa: E402


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--reps", type=int, default=200)
    args = ap.parse_args()
    sc = scenes.c2_sheet(n=args.n)
    s = HipSolver()
    s.build(sc)
    _, _, blocks = s.debug_assemble(1 / 60, np.zeros((sc.n, 3), np.float32))
    st = s.structure()
    x = np.random.default_rng(1).uniform(-1, 1, (sc.n, 3)).astype(np.float32)
    y, us = bsr_spmv(st["row_start"], st["col_index"], blocks, None, x, args.reps)
    ys = s.debug_spmv(x)
    # FP64 product of the same blocks on the host, the reference both device products are checked against.
    rs = st["row_start"].astype(np.int64)
    rows = np.repeat(np.arange(sc.n), np.diff(rs))
    y64 = np.zeros((sc.n, 3))
    np.add.at(y64, rows, np.einsum("kij,kj->ki", blocks.astype(np.float64), x[st["col_index"]].astype(np.float64)))
    norm = np.linalg.norm(y64)
    nb = len(st["col_index"])
    alg = 40.0 * nb + 4.0 * (sc.n + 1) + 24.0 * sc.n
    print(json.dumps({"kernel": "kSpmv16", "particles": sc.n, "blocks": nb, "reps": args.reps,
                      "mean_us_events": us, "algorithmic_bytes": alg, "achieved_gbs_events": alg / us / 1e3,
                      "relative_error_bsr_spmv": float(np.linalg.norm(y - y64) / norm),
                      "relative_error_solver_spmv": float(np.linalg.norm(ys - y64) / norm),
                      "max_row_error_solver_spmv": float(np.max(np.abs(ys - y64)) / np.max(np.abs(y64)))}),
          flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
