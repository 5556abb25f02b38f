"""Writes the FP64-oracle fixture of the 65,536-particle sheet drop: positions after 10 and 20 implicit steps
(dt 1/60 s, 1 substep, Newton <= 5, PCG to 1e-6 with at most 2,000 iterations: the benchmark's step) and the
oracle's own spread at those steps when its start positions move by up to 1e-9 m (uniform per component, seed 1).

Run from the repository root: python tests/golden/make_c2_drop.py   (about two minutes on 8 host threads)
Output: tests/golden/c2_drop_oracle.npz
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(REPO, "drape-2_amd"))
sys.path.insert(0, os.path.join(REPO, "oracle"))

import pyoracle  # noqa: E402
from drape_amd import scenes  # noqa: E402

CHECK_STEPS = (10, 20)
NUDGE = 1e-9


def run(sc, nudge):
    s = pyoracle.OracleSolver()
    s.build(sc)
    if nudge > 0:
        x = s.read_positions()
        rng = np.random.default_rng(1)
        s.write_state(x=x + nudge * rng.uniform(-1.0, 1.0, size=x.shape))
    out = {}
    for k in range(1, max(CHECK_STEPS) + 1):
        s.step(1 / 60, 1, 5)
        if k in CHECK_STEPS:
            out[k] = s.read_positions().astype(np.float64)
            print(f"  step {k}: {s.counters()}", flush=True)
    return out


def main():
    pyoracle.set_threads(os.cpu_count() or 1)
    sc = scenes.c2_sheet()
    t0 = time.time()
    base = run(sc, 0.0)
    nudged = run(sc, NUDGE)
    arrays = {}
    for k in CHECK_STEPS:
        d = np.linalg.norm(base[k] - nudged[k], axis=1)
        arrays[f"x{k}"] = base[k].astype(np.float32)
        arrays[f"spread{k}"] = np.array([d.max(), np.sqrt(np.mean(d * d))])
        print(f"step {k}: oracle self-spread max {d.max() * 1e3:.4f} mm rms {np.sqrt(np.mean(d * d)) * 1e3:.4f} mm")
    np.savez_compressed(os.path.join(HERE, "c2_drop_oracle.npz"), steps=np.array(CHECK_STEPS), nudge=NUDGE, **arrays)
    print(f"done in {time.time() - t0:.0f} s")


if __name__ == "__main__":
    main()
