"""Benchmark: implicit steps/s of the Newton–PCG cloth step on one MI355X.

Workload (implicit steps/s at 65k particles; the learned warm start against the plain guess in the line's
`warm_start_ab` block): the 256x256-vertex sheet (65,536 particles,
130,050 stretch triangles, 194,565 hinges, 4 mm spacing) dropped from y = 0.4 m onto a sphere (r 0.25 m) plus a
ground plane (y = -0.3 m), one implicit-Euler step per frame (dt = 1/60 s, 1 substep, Newton <= 5, block-Jacobi
PCG to relative residual 1e-6), FP32 assembly/SpMV/PCG with FP64 reductions. A "step" is one Solver::step call;
the timed region holds K consecutive steps of the drop after W warm-up steps, with every input resident in HBM.

Printed as one JSON line with the measured roofline fraction of the dominant kernel (the fused PCG SpMV, timed
with HIP events on the solver's stream inside the timed region) and the FP64 CPU oracle timed on the host as
the CPU baseline.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "drape-2_amd"))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
# Per PCG iteration on the 256x256 sheet, from rocprofv3 runs of the same workload: the kernel's mean time in the
# kernel trace, and its HBM traffic (FETCH_SIZE and WRITE_SIZE scaled by a plain copy's readings, x2.0 and x1.0).
# kPcgStep (one launch per iteration): profiles/kPcgStep_sheet65k_kernel_trace.md (launches that ran an iteration).
# kPcgSolve (the persistent solve, one launch per solve): profiles/kPcgSolve_sheet65k_kernel_trace.md (kernel time and
# counter sums over the run divided by its PCG iterations).
KERNEL_FIGURES = {
    (1, 8, 0): {"kernel": "kPcgStep (one-kernel PCG iteration, one launch per iteration: update, convergence test, "
                          "SELL-64 block SpMV with the search direction rebuilt per gathered column, contact filter, "
                          "FP64 partials)",
                "trace_us": 9.62, "traffic": 35.10e6},
    (1, 8, 2): {"kernel": "kPcgSolve (the persistent solve: one launch runs a whole PCG solve of the same one-kernel "
                          "iteration, the matrix and each row's private vectors held in registers, a barrier per "
                          "scene per iteration)",
                "trace_us": 6.301, "traffic": 5.18e6},
}


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def cpu_baseline(make_scene, snapshot: bytes, runs: int = 3) -> tuple:
    """The FP64 oracle (this project's restatement of the reference's single-threaded CpuSolver) on one host
    thread: one implicit step from the GPU's own state at the start of the timed window (its snapshot restored into
    the oracle), `runs` times; the median. Returns the baseline and the oracle's positions after that step."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import pyoracle
    pyoracle.set_threads(1)
    s = pyoracle.OracleSolver()
    s.build(make_scene())
    secs = []
    for _ in range(runs):
        s.restore(snapshot)
        c0 = s.counters()
        t0 = time.perf_counter()
        s.step(1 / 60, 1, 5)
        secs.append(time.perf_counter() - t0)
    c = s.counters()
    med = float(np.median(secs))
    return ({"value": 1.0 / med, "unit": "implicit steps/s", "cores": 1, "kind": "port",
             "sample": f"one implicit step from the GPU's state at the start of the timed window, restored into the "
                       f"FP64 oracle, 1 thread, median of {runs} runs ({', '.join(f'{x:.2f}' for x in secs)} s; "
                       f"{c['newton_iterations'] - c0['newton_iterations']} Newton, "
                       f"{c['pcg_iterations'] - c0['pcg_iterations']} PCG iterations per run), {cpu_model()}, "
                       f"nproc {os.cpu_count()}"},
            s.read_positions().astype(np.float64))


def survey_pcg_bytes(z: dict) -> float:
    """Algorithmic bytes of one PCG iteration as the scope table defines them: the block SpMV
    nnzb (36 + 4) + 4 (n + 1) + 12 n + 12 n, plus 14 vector passes of 12 n."""
    n, nnzb = z["n_particles"], z["n_blocks"]
    return nnzb * 40.0 + 4.0 * (n + 1) + 24.0 * n + 14.0 * 12.0 * n


GOLDEN = os.path.join(REPO, "tests", "golden", "c2_drop_oracle.npz")


WEIGHTS = os.path.join(REPO, "drape-2_amd", "drape_amd", "weights", "warmstart.npy")


def warm_start_ab(make_scene, warmup: int, steps: int, net_path: str, caps=(3, 4, 5), ref_iterations: int = 40,
                  sample_every: int = 2, fixture=None) -> dict:
    """Learned warm start against the previous-velocity guess at matched error. Each method steps the drop with
    Newton capped at each of `caps`; every `sample_every`-th timed step the converged plain solve (Newton <=
    ref_iterations, same PCG) from that method's own state before the step judges the step's error (max and RMS
    particle distance, mm). Reports steps/s over the timed steps (the reference solves excluded), iteration counts,
    how often the energy check kept the learned guess, and for every (learned cap, plain cap) pair the speed ratio
    and the ratios of the three error measures (below 1: the learned run is the more accurate). With the FP64-oracle
    fixture of the drop (positions at given steps from rest, plain guess, Newton <= 5), each run's own trajectory
    from rest is also compared with the oracle at those steps, beside the oracle's self-spread there."""
    from drape_amd import HipSolver, learn
    net = learn.WarmStartNet()
    net.load(net_path)
    sc = make_scene()
    ref = HipSolver()
    ref.build(sc)
    runs = []
    for learned in (False, True):
        for k in caps:
            s = HipSolver()
            s.build(sc)
            if learned:
                s.set_warm_start(net)
            errs, rms, t = [], [], 0.0
            oracle = {}
            for i in range(warmup + steps):
                check = i >= warmup and (i - warmup) % sample_every == 0
                if check:
                    snap = s.snapshot()
                if i == warmup:
                    s.reset_counters()
                t0 = time.perf_counter()
                s.step(1 / 60, 1, k)
                if i >= warmup:
                    t += time.perf_counter() - t0
                if fixture is not None and i + 1 in [int(x) for x in fixture["steps"]]:
                    d = np.linalg.norm(s.read_positions().astype(np.float64) - fixture[f"x{i + 1}"].astype(np.float64),
                                       axis=1)
                    oracle[str(i + 1)] = {"max_mm": float(d.max()) * 1e3,
                                          "rms_mm": float(np.sqrt(np.mean(d * d))) * 1e3,
                                          "oracle_self_spread_mm": float(fixture[f"spread{i + 1}"][0]) * 1e3}
                if check:
                    ref.restore(snap)
                    ref.step(1 / 60, 1, ref_iterations)
                    d = np.linalg.norm(s.read_positions().astype(np.float64) - ref.read_positions(), axis=1)
                    errs.append(float(d.max()) * 1e3)
                    rms.append(float(np.sqrt(np.mean(d * d))) * 1e3)
            c = s.counters()
            runs.append({"guess": "learned" if learned else "previous velocity", "newton_cap": k,
                         "steps_per_s": steps / t, "newton_per_step": c["newton_iterations"] / steps,
                         "pcg_per_step": c["pcg_iterations"] / steps,
                         "mean_rms_error_mm": float(np.mean(rms)), "mean_max_error_mm": float(np.mean(errs)),
                         "max_step_error_mm": max(errs), "median_step_error_mm": float(np.median(errs)),
                         "steps_within_0p1mm": int(sum(e <= 0.1 for e in errs)), "checked_steps": len(errs),
                         "learned_guess_kept": c["warm_start_accepted"],
                         "learned_guess_rejected": c["warm_start_fallbacks"],
                         "fp64_oracle_from_rest": oracle})
    plain = [r for r in runs if r["guess"] != "learned"]
    pairs = []
    for r in runs:
        if r["guess"] != "learned":
            continue
        for p in plain:
            pairs.append({"learned_cap": r["newton_cap"], "plain_cap": p["newton_cap"],
                          "speedup": r["steps_per_s"] / p["steps_per_s"],
                          "rms_error_ratio": r["mean_rms_error_mm"] / p["mean_rms_error_mm"],
                          "mean_max_error_ratio": r["mean_max_error_mm"] / p["mean_max_error_mm"],
                          "worst_step_error_ratio": r["max_step_error_mm"] / p["max_step_error_mm"]})
    return {"error_reference": f"converged plain solve (Newton <= {ref_iterations}) from the method's own state "
                               f"before the step, every {sample_every} timed steps; fp64_oracle_from_rest: the run's "
                               f"own trajectory from rest against the FP64 oracle's (plain guess, Newton <= 5)",
            "weights": os.path.relpath(net_path, REPO), "runs": runs, "learned_vs_plain": pairs}


def algorithmic_step_bytes(s, counters: dict, steps: int, pcg_launch_bytes: float) -> float:
    """Algorithmic bytes of one implicit step of the timed run: per Newton iteration the assembly (read x start,
    u, each term's constants and the contact state; write the stored blocks and the gradient) and the
    preconditioner; per PCG iteration its algorithmic bytes; per line-search trial an energy-only evaluation."""
    z = s.structure_sizes()
    n, S, H, K = z["n_particles"], z["n_stretch"], z["n_hinges"], z["n_seams"]
    upper = 0.5 * (z["n_blocks"] + n)
    terms = 24.0 * n + 36.0 * S + 24.0 * H + 16.0 * K + 32.0 * n
    assembly = terms + 12.0 * n + 36.0 * upper
    per_step = lambda key: counters[key] / steps  # noqa: E731
    return (per_step("newton_iterations") * (assembly + 48.0 * n) + per_step("pcg_iterations") * pcg_launch_bytes
            + per_step("line_search_trials") * terms)


# The reference's published tee drapes to rest (men's tee, size M, Apple M1): particles, CPU wall s and frames,
# Metal wall s and frames (docs/superpowers/specs/2026-10-25-drape-phase-b-mac-app-design.md:186-194).
TEE_PUBLISHED = {"draft": (2647, 10.00, 418, 13.13, 507), "standard": (10342, 38.59, 397, 19.21, 375),
                 "fine": (40046, 397.50, 480, 107.02, 351)}
TEE_CPU_FRAMES = {"draft": 120, "standard": 30, "fine": 6}


def tee_bench(quality: str, no_cpu: bool) -> dict:
    """The men's tee (size M) on the male mannequin draped to rest on the GPU: wall time from the first frame to
    the settled one, after the scene is built (mesh, placement, GPU bake of the body field, solver build)."""
    from drape_amd import HipSolver, bake_collision_field, tee
    m = tee.load_mannequin("male")
    t0 = time.perf_counter()
    field = bake_collision_field([(m.vertices, m.triangles)])
    bake_s = time.perf_counter() - t0
    ts = tee.tee_scene("male", quality, field=field)
    s = HipSolver()
    s.build(ts.scene)
    opts = tee.options_for(quality)
    t0 = time.perf_counter()
    r = tee.drape_to_rest(s, ["tee"], opts)
    wall = time.perf_counter() - t0
    st = s.stats()
    c = s.counters()
    pub = TEE_PUBLISHED[quality]
    line = {
        "metric": f"tee drape to rest ({quality})", "value": wall, "unit": "s", "higher_is_better": False,
        "n_gpus": 1, "dtype": "f32 (FP64 reductions)", "data": "synthetic pattern, reference mannequin asset",
        "vs_baseline": None,
        "config": {"workload": f"men's test tee size M on the male mannequin, {quality} ({ts.scene.n} particles)",
                   "particles": ts.scene.n, "substeps": opts.substeps, "newton": opts.iterations,
                   "pcg": "<= 50 iterations to 1e-2", "settled": r.settled, "failed": r.failed,
                   "frames": r.frames_run, "simulated_s": r.simulated_seconds, "retries": r.retries,
                   "assembly_s": r.assembly_seconds, "assembly_end_seam_gap_mm": r.assembly_end_seam_gap * 1e3,
                   "max_penetration_mm": st["max_body_penetration"] * 1e3,
                   "hem_error_mm": (tee.hem_height(ts, s.read_positions()) - tee.hem_target(ts)) * 1e3,
                   "frames_per_s": r.frames_run / wall, "body_field_bake_s": bake_s,
                   "newton_iterations": c["newton_iterations"], "pcg_iterations": c["pcg_iterations"]},
        "published_m1": {"particles": pub[0], "cpu_s": pub[1], "cpu_frames": pub[2], "metal_s": pub[3],
                         "metal_frames": pub[4]},
    }
    if not no_cpu:
        sys.path.insert(0, os.path.join(REPO, "oracle"))
        import pyoracle
        pyoracle.set_threads(1)
        o = pyoracle.OracleSolver()
        o.build(ts.scene)
        o.set_phase("tee", tee.ASSEMBLING)
        n = TEE_CPU_FRAMES[quality]
        t0 = time.perf_counter()
        for _ in range(n):
            o.step(opts.frame_dt, opts.substeps, opts.iterations)
        dt = time.perf_counter() - t0
        line["cpu_baseline"] = {"value": n / dt, "unit": "frames/s", "cores": 1, "kind": "port",
                                "sample": f"the first {n} frames of the same drape (assembling), FP64 oracle, "
                                          f"1 thread, {cpu_model()}"}
    return line


def batch_bench(steps: int, warmup: int, sample_every: int, shape: tuple, graph: bool = True) -> dict:
    """The 64-sheet batch (log-uniform 2k-40k particles per sheet, each on its own sphere or capsule plus ground,
    PCG to 1e-6 with the cap at 2,000): one solver, one launch per kernel for all 64 segments."""
    from drape_amd import HipSolver, scenes
    batch = scenes.c4_batch()
    n = sum(sc.n for sc in batch)
    s = HipSolver(graph=graph, pcg_shape=shape)
    s.build_batch(batch)
    s.set_spmv_sampling(sample_every)
    for _ in range(warmup):
        s.step(1 / 60, 1, 5)
    s.reset_counters()
    s.set_spmv_sampling(sample_every)
    t0 = time.perf_counter()
    for _ in range(steps):
        s.step(1 / 60, 1, 5)
    elapsed = time.perf_counter() - t0
    c = s.counters()
    timing = s.spmv_timing()
    stats = s.segment_stats()
    achieved = timing["bytes_per_launch"] / (timing["mean_us"] * 1e-6) / 1e9 if timing["mean_us"] > 0 else None
    w, k, h = s.pcg_shape()
    # The scope table's bytes of an iteration (a full block SpMV and 14 vector passes) over the layout's, for the
    # whole batch: the same ratio holds segment by segment to within a percent (every segment is a grid sheet).
    z = s.structure_sizes()
    nz, npart = float(z["n_blocks"]), float(z["n_particles"])
    scope_ratio = survey_pcg_bytes(z) / (40.0 * 0.5 * (nz + npart) + 8.0 * 0.5 * (nz - npart) + 192.0 * npart)
    return {
        "metric": "batched implicit steps/sec, 64 uneven sheets", "value": steps / elapsed,
        "unit": "implicit steps/s", "n_gpus": 1, "steps": steps, "warmup": warmup,
        "ms_per_step": 1000.0 * elapsed / steps, "higher_is_better": True, "scaling": "weak",
        "dtype": "f32 (FP64 reductions)", "data": "synthetic", "vs_baseline": None,
        "config": {"workload": f"64 sheets, {n} particles (2k-40k each), dropped on spheres/capsules + ground",
                   "particles": n, "particle_steps_per_s": n * steps / elapsed,
                   "pcg_launch_shape": f"{w}x{k}{('', 'h', 'p')[h]}",
                   "newton_per_step": c["newton_iterations"] / steps, "pcg_per_step": c["pcg_iterations"] / steps,
                   "non_finite_segments": sum(bool(x["has_non_finite"]) for x in stats),
                   "max_penetration_mm": 1e3 * max(x["max_body_penetration"] for x in stats)},
        "roofline": {"kernel": ("kPcgSolve (the persistent solve: each segment's tiles resident together, the "
                                "segments queued through the CUs; bytes count each segment's own iterations)"
                                if h == 2 else
                                "kPcgStep (one launch covers all 64 segments; bytes count the segments still "
                                "solving at each timed launch)"),
                     "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": (achieved / HBM_PEAK_GBS) if achieved else None, "traffic": None,
                     "bytes_per_launch": timing["bytes_per_launch"], "mean_launch_us": timing["mean_us"],
                     "timed_launches": timing["launches"],
                     "bytes_definition": "the symmetric layout's bytes: 40 B per block on or above the diagonal, 8 B "
                                         "per slot below it, 192 B per particle",
                     "frac_scope_bytes": (achieved * scope_ratio / HBM_PEAK_GBS) if achieved else None,
                     "scope_bytes_ratio": scope_ratio},
    }


def tube_bench(steps: int, shape: tuple, graph: bool = True) -> dict:
    """The seamed tube (two 200 x 300 panels sewn into a tube, 120,000 particles) falling onto the capsule
    mannequin and settling: mean time per implicit step over `steps` steps from the start, and the last 100."""
    from drape_amd import HipSolver, scenes
    sc = scenes.c3_tube()
    s = HipSolver(graph=graph, pcg_shape=shape)
    s.build(sc)
    times = []
    for _ in range(steps):
        t0 = time.perf_counter()
        s.step(1 / 60, 1, 5)
        times.append(time.perf_counter() - t0)
    st = s.stats()
    c = s.counters()
    w, k, h = s.pcg_shape()
    tail = times[-100:] if len(times) >= 100 else times
    return {
        "metric": "seamed tube, ms per implicit step", "value": 1000.0 * sum(times) / len(times), "unit": "ms",
        "higher_is_better": False, "n_gpus": 1, "steps": steps, "dtype": "f32 (FP64 reductions)",
        "data": "synthetic", "vs_baseline": None,
        "config": {"workload": f"tube of two 200x300 panels ({sc.n} particles), seams k = 1e3 N/m, on the capsule "
                               "mannequin, PCG to 1e-6 (cap 2,000), Newton <= 5",
                   "last_100_steps_ms": 1000.0 * sum(tail) / len(tail),
                   "pcg_launch_shape": f"{w}x{k}{('', 'h', 'p')[h]}",
                   "newton_per_step": c["newton_iterations"] / steps, "pcg_per_step": c["pcg_iterations"] / steps,
                   "final_stats": st},
    }


def newton_curve(net_path: str, steps: int = 60, every: int = 3, caps=(1, 2, 3, 5, 8, 12, 20, 30),
                 ref_iterations: int = 40) -> dict:
    """How far a step is from the converged solve after k Newton iterations. The sheet drop runs with the plain
    guess (Newton <= 5); every `every`-th step from step 3, its state is restored into a plain-guess and a learned-
    guess solver, each steps once with Newton capped at every k of `caps`, and the max particle distance (mm) to the
    Newton <= `ref_iterations` plain solve from the same state is recorded. Restoring a snapshot clears the realised
    dv the network reads, so the learned guess here sees dv_prev = 0."""
    from drape_amd import HipSolver, learn, scenes
    net = learn.WarmStartNet()
    net.load(net_path)
    sc = scenes.c2_sheet()
    run, ref, plain, learned = (HipSolver() for _ in range(4))
    for s in (run, ref, plain, learned):
        s.build(sc)
    learned.set_warm_start(net)
    rows = []
    for i in range(steps + 2):
        if i >= 2 and (i - 2) % every == 0:
            snap = run.snapshot()
            ref.restore(snap)
            ref.reset_counters()
            ref.step(1 / 60, 1, ref_iterations)
            x_ref = ref.read_positions().astype(np.float64)
            row = {"step": i + 1, "ref_newton": ref.counters()["newton_iterations"]}
            for name, s in (("plain", plain), ("learned", learned)):
                errs = []
                for k in caps:
                    s.restore(snap)
                    s.reset_counters()
                    s.step(1 / 60, 1, k)
                    errs.append(round(float(np.linalg.norm(s.read_positions() - x_ref, axis=1).max()) * 1e3, 4))
                row[name] = errs
                if name == "learned":
                    row["learned_guess_kept"] = bool(s.counters()["warm_start_accepted"])
            rows.append(row)
            print(json.dumps(row), file=sys.stderr, flush=True)
        run.step(1 / 60, 1, 5)
    summary = {}
    for name in ("plain", "learned"):
        summary[name] = {str(k): {"within_0p1mm": sum(r[name][j] <= 0.1 for r in rows),
                                  "median_max_mm": float(np.median([r[name][j] for r in rows]))}
                         for j, k in enumerate(caps)}
    return {"metric": "max distance to the converged solve after k Newton iterations, sheet drop",
            "unit": "mm", "caps": list(caps), "reference_newton_cap": ref_iterations, "checked_steps": len(rows),
            "summary": summary, "steps": rows}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=256, help="sheet vertices per side (256 = the 65,536-particle sheet)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--sample-every", type=int, default=8,
                    help="event-time PCG iteration launches (graph path: one per poll; eager: every k-th)")
    ap.add_argument("--pcg-variant", type=int, default=1, help="1 single-kernel PCG (default), 0 two-kernel")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--no-refinement", action="store_true", help="turn the FP64 true-residual refinement off")
    ap.add_argument("--no-warm-start-ab", action="store_true",
                    help="skip the learned-vs-plain warm start comparison")
    ap.add_argument("--warm-start-weights", default=WEIGHTS)
    ap.add_argument("--learned", action="store_true",
                    help="seed Newton with the learned warm start in the timed run (default: the plain guess)")
    ap.add_argument("--newton", type=int, default=5, help="Newton iteration cap of the timed run")
    ap.add_argument("--pcg-shape", default="0x0",
                    help="launch shape of the one-kernel PCG iteration, wavefronts x slots: 1x8, 1x4, 2x8, 4x8, "
                         "or 1x8h / 1x4h (column window staged in LDS), or 1x8p (the persistent solve: one kernel "
                         "launch per PCG solve); 0x0 = tune during warm-up")
    ap.add_argument("--tee", choices=("draft", "standard", "fine"),
                    help="instead of the sheet benchmark: drape the reference's tee to rest at this quality")
    ap.add_argument("--tube", action="store_true",
                    help="instead of the sheet benchmark: the seamed tube settling on the mannequin (--steps steps)")
    ap.add_argument("--batch", action="store_true",
                    help="instead of the sheet benchmark: the 64-sheet batch (one solver, one launch per kernel)")
    ap.add_argument("--newton-curve", action="store_true",
                    help="instead of the sheet benchmark: distance to the converged solve after k Newton iterations")
    args = ap.parse_args()
    spec = args.pcg_shape.lower()
    shape = tuple(int(t) for t in spec.rstrip("hp").split("x")) + (2 if spec.endswith("p") else int(spec.endswith("h")),)
    if args.tube:
        print(json.dumps(tube_bench(args.steps, shape, not args.no_graph)), flush=True)
        return 0
    if args.batch:
        print(json.dumps(batch_bench(args.steps, args.warmup, args.sample_every, shape, not args.no_graph)),
              flush=True)
        return 0
    if args.tee:
        print(json.dumps(tee_bench(args.tee, args.no_cpu_baseline)), flush=True)
        return 0
    if args.newton_curve:
        print(json.dumps(newton_curve(args.warm_start_weights)), flush=True)
        return 0
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported by this build "
              f"(the batched path shards by scene, see DESIGN.md)", file=sys.stderr)
        return 2

    from drape_amd import HipSolver, scenes

    def make_scene():
        return scenes.c2_sheet(n=args.n)

    sc = make_scene()
    s = HipSolver(graph=not args.no_graph, pcg_variant=args.pcg_variant, pcg_shape=shape)
    s.build(sc)
    net = None
    if args.learned:
        from drape_amd import learn
        net = learn.WarmStartNet()
        net.load(args.warm_start_weights)
        s.set_warm_start(net)
    if args.no_refinement:
        s.set_refinement(0)
    s.set_spmv_sampling(args.sample_every)
    for _ in range(args.warmup):
        s.step(1 / 60, 1, args.newton)
    window_start = s.snapshot()
    s.reset_counters()
    s.set_spmv_sampling(args.sample_every)
    fixture = np.load(GOLDEN) if args.n == 256 and os.path.exists(GOLDEN) else None
    checkpoints, elapsed = [], 0.0
    for i in range(args.steps):
        t0 = time.perf_counter()
        s.step(1 / 60, 1, args.newton)  # returns after the device finished the step
        elapsed += time.perf_counter() - t0
        k = args.warmup + i + 1  # steps from rest
        if fixture is not None and k in [int(x) for x in fixture["steps"]]:
            d = np.linalg.norm(s.read_positions().astype(np.float64) - fixture[f"x{k}"].astype(np.float64), axis=1)
            checkpoints.append({"step": k, "max_mm": float(d.max()) * 1e3,
                                "rms_mm": float(np.sqrt(np.mean(d * d))) * 1e3,
                                "oracle_self_spread_mm": float(fixture[f"spread{k}"][0]) * 1e3})
    c = s.counters()
    rep = s.pcg_report()
    timing = s.spmv_timing()
    st = s.stats()
    value = args.steps / elapsed
    z = s.structure_sizes()
    pcg_bytes = survey_pcg_bytes(z)
    step_bytes = algorithmic_step_bytes(s, c, args.steps, pcg_bytes)
    per_iter_us = timing["mean_us"]
    shape_in_use = tuple(s.pcg_shape())
    persistent = shape_in_use[2] == 2
    fig = KERNEL_FIGURES.get(shape_in_use) if args.pcg_variant == 1 and args.n == 256 else None
    achieved = pcg_bytes / (per_iter_us * 1e-6) / 1e9 if per_iter_us > 0 else None
    line = {
        "metric": "implicit steps/sec @65k particles; learned-warm-start vs plain at 0.1 mm error",
        "value": value,
        "unit": "implicit steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1000.0 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (FP64 reductions)",
        "data": "synthetic",
        "config": {
            "workload": f"sheet {args.n}x{args.n} ({sc.n} particles) dropped from y=0.4 m on sphere r=0.25 m + "
                        f"ground y=-0.3 m; steps {args.warmup + 1}..{args.warmup + args.steps} of the drop",
            "particles": sc.n,
            "dt": 1 / 60, "substeps": 1, "newton_max": args.newton, "pcg_tolerance": 1e-6, "pcg_cap": 2000,
            "solver": ("learned warm start (u = y + h d, network " + os.path.relpath(args.warm_start_weights, REPO)
                       + ")") if args.learned else "plain warm start (u = h v)",
            "pcg_form": "one-kernel" if args.pcg_variant == 1 else "two-kernel",
            "pcg_launch_shape": "%dx%d%s" % (s.pcg_shape()[0], s.pcg_shape()[1], ("", "h", "p")[s.pcg_shape()[2]])
                                if args.pcg_variant == 1 else None,
            "fp64_true_residual_refinement": not args.no_refinement,
            "graph": not args.no_graph,
            "newton_per_step": c["newton_iterations"] / args.steps,
            "pcg_per_step": c["pcg_iterations"] / args.steps,
            "line_search_trials_per_step": c["line_search_trials"] / args.steps,
            "final_stats": st,
        },
        "pcg": {"solves": rep["solves"], "exit_tolerance": rep["exit_tolerance"], "exit_cap": rep["exit_cap"],
                "exit_curvature": rep["exit_curvature"], "refinement_restarts": rep["restarts"],
                "true_rel_residual_mean": rep["true_rel_mean"], "true_rel_residual_max": rep["true_rel_max"],
                "true_rel_residual_max_tolerance_exits": rep["true_rel_max_converged"],
                "true_rel_over_tol_or_fp32_floor_max": rep["true_rel_excess_max"]},
        "roofline": {
            "kernel": (fig or KERNEL_FIGURES[(1, 8, 0)])["kernel"] if args.pcg_variant == 1 else
                      "kPcgSpmv (fused p = z + beta p, SELL-64 block SpMV, contact filter, FP64 p.Ap partials)",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
            "traffic": fig["traffic"] if fig else None,
            "bytes_per_launch": pcg_bytes,
            "bytes_definition": "scope-table algorithmic bytes of one PCG iteration: full block SpMV nnzb (36 + 4) "
                                "+ 4 (n + 1) + 24 n, plus 14 vector passes of 12 n",
            "layout_bytes_per_launch": timing["bytes_per_launch"],
            "per_iteration_us": per_iter_us,
            "timing": ("HIP events around each persistent solve launch of the timed steps, divided by the PCG "
                       "iterations it ran" if persistent else
                       "HIP events around whole replayed 64-launch chunks of the timed steps, counted when the solve "
                       "was still running after the chunk (every launch did an iteration); the gaps between launches "
                       "are included"),
            "timed_iterations": timing["launches"],
            "kernel_trace_us": fig["trace_us"] if fig else None,
            "kernel_trace_frac": (pcg_bytes / (fig["trace_us"] * 1e-6) / 1e9 / HBM_PEAK_GBS) if fig else None,
        },
    }
    if persistent:
        line["roofline"]["on_chip"] = (
            "the persistent solve keeps the matrix and each row's private vectors in registers for the whole solve, "
            "so an iteration moves only the gathered window, the rows' new (zr, t, p) and the partial sums: `traffic` "
            "(HBM bytes per iteration, PMC) is far below the algorithmic bytes, and the iteration's time is set by its "
            "per-scene barrier and exchange round trips, not by HBM bandwidth; `frac` compares that time with the "
            "algorithmic bytes of a matrix-streaming iteration at the HBM peak")
    line["step_roofline"] = {
        "bytes_per_step": step_bytes,
        "achieved": step_bytes * value / 1e9,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": step_bytes * value / 1e9 / HBM_PEAK_GBS,
        "terms": "assembly 24n + 36 S + 24 H + 16 K + 44n + 36 U per Newton iteration, preconditioner 48n per Newton "
                 "iteration, one PCG iteration's scope-table bytes per PCG iteration, 24n + 36 S + 24 H + 16 K + 32n "
                 "per line-search "
                 "trial (n particles, S stretch, H hinge, K seam terms, U stored blocks)",
    }
    accuracy = {"reference": "FP64 oracle (this project's restatement of CpuSolver), same scene and step settings",
                "checkpoints_from_rest": checkpoints}
    if not args.no_cpu_baseline:
        base, x_oracle = cpu_baseline(make_scene, window_start)
        line["cpu_baseline"] = base
        s.restore(window_start)
        s.step(1 / 60, 1, 5)
        d = np.linalg.norm(s.read_positions().astype(np.float64) - x_oracle, axis=1)
        accuracy["one_step_from_window_start"] = {"step": args.warmup + 1, "max_mm": float(d.max()) * 1e3,
                                                  "rms_mm": float(np.sqrt(np.mean(d * d))) * 1e3}
    line["accuracy"] = accuracy
    if not args.no_warm_start_ab and os.path.exists(args.warm_start_weights):
        line["warm_start_ab"] = warm_start_ab(make_scene, args.warmup, args.steps, args.warm_start_weights,
                                              fixture=fixture)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
