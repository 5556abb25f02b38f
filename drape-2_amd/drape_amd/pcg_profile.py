"""Driver for profiling the PCG kernels.

Steps a scene `--steps` times with a fixed PCG launch shape and prints one JSON line with the solver's counters:
the 256x256 sheet drop (the benchmark scene, default), the 120,000-particle seamed tube on the capsule body
(`--scene tube`) or the 64 uneven sheets batched (`--scene batch`). Under `rocprofv3 --kernel-trace --stats` (or
`--pmc`), a kernel's total time (or counter sum) divided by the PCG iterations printed here gives its per-iteration
figure; this matters for the persistent solve (shape 1x8p), whose one launch runs a whole solve. For the batch the
printed iterations are summed over its scenes.

    python -m drape_amd.pcg_profile --shape 1x8p --steps 10
"""
from __future__ import annotations

import argparse
import json
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

from drape_amd import HipSolver, scenes  # noqa: E402


def parse_shape(spec: str) -> tuple:
    """'1x8' -> (1, 8, 0); '1x8h' -> (1, 8, 1); '1x8p' -> (1, 8, 2) (the persistent solve)."""
    spec = spec.lower()
    code = 2 if spec.endswith("p") else 1 if spec.endswith("h") else 0
    w, k = (int(t) for t in spec.rstrip("hp").split("x"))
    return (w, k, code)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--shape", default="1x8p")
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--scene", choices=("sheet", "tube", "batch"), default="sheet")
    args = ap.parse_args()
    s = HipSolver(graph=not args.no_graph, pcg_shape=parse_shape(args.shape))
    if args.scene == "batch":
        batch = scenes.c4_batch()
        s.build_batch(batch)
        label = f"64 sheets batched, {sum(x.n for x in batch)} particles"
    else:
        sc = scenes.c3_tube() if args.scene == "tube" else scenes.c2_sheet(n=args.n)
        s.build(sc)
        label = (f"seamed tube, {sc.n} particles" if args.scene == "tube"
                 else f"sheet {args.n}x{args.n}, {sc.n} particles")
    for _ in range(args.steps):
        s.step(1 / 60, 1, 5)
    c = s.counters()
    rep = s.pcg_report()
    print(json.dumps({"scene": label, "steps": args.steps,
                      "shape": args.shape, "shape_in_use": list(s.pcg_shape()), "counters": c,
                      "solves": rep["solves"], "refinement_restarts": rep["restarts"],
                      "persistent_aborts": rep["persistent_aborts"]}), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
