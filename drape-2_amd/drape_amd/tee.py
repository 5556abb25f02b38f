"""The reference's tee scene end to end on the MI355X solver: mannequin, tee pattern, simulation mesh, placement,
collision field, and the assemble-then-settle loop.

* `load_mannequin` reads a body's triangle mesh and region cylinders (libs/avatar/src/mannequin.cpp:54-118,
  obj_reader.cpp:23-58) from the assets shipped in `assets/mannequins`.
* `tee_scene` follows makeTeeScene (libs/session/src/tee_scene.cpp:24-40): the men's cut for the male body and
  the women's for the female, size M, particle distance by quality (sim_mesh_builder.cpp:14-21), placement on
  the body-region cylinders, jersey for the body pieces and rib for the neck rib (test_fabrics.cpp:5-6), and the
  body's collision field baked on the GPU (cell 4 mm, padding 5 cm, band 2 cm: collision_bake.hpp:10-14).
* `drape_to_rest` is the reference's DrapeSession state machine (libs/sim/src/drape_session.cpp:35-95) with
  DrapeOptions' defaults (drape_runner.hpp:14-26): seams close in the assembling phase until the largest seam
  gap is below 1 mm (at most 3 s), then the garment settles until its rms speed stays below 2 mm/s for 60
  frames; non-finite values or 60 frames of rising kinetic energy roll back to the last snapshot (every 30
  frames) with the substeps doubled, at most 3 times; 20 s of simulated time at most. The statistics come from
  the solver's device reduction.
* Substeps x Newton iterations per frame by quality: 1 x 2 (draft, standard), 2 x 2 (fine)
  (cpu_solver.cpp:16-32).
"""
from __future__ import annotations

import dataclasses
import json
import math
import os

import numpy as np

from . import garment as G
from . import scenes

ASSETS = os.path.join(os.path.dirname(__file__), "assets", "mannequins")

LANDMARKS = ("neck_base", "hps_left", "hps_right", "shoulder_left", "shoulder_right", "chest", "waist", "hip",
             "crotch", "shoulder_joint_left", "shoulder_joint_right", "elbow_left", "elbow_right", "wrist_left",
             "wrist_right", "hip_joint_left", "hip_joint_right", "ankle_left", "ankle_right")
GIRTHS = ("chest", "waist", "hip", "neck")


def test_rib() -> dict:
    """The reference's rib test fabric (libs/core/src/test_fabrics.cpp:6)."""
    return scenes.fabric(300, 100, 80, 0.8, 1.0, 3e-6, 2e-6, 0.280, 0.0010, 0.3, 0.3)


@dataclasses.dataclass
class Mannequin:
    vertices: np.ndarray   # (n, 3) float64, metres
    triangles: np.ndarray  # (T, 3) int64
    height: float
    girths: dict
    landmarks: dict
    regions: dict          # region name -> garment.RegionCylinder


def read_obj_triangles(path: str):
    """Vertices and triangles of an OBJ file ('v x y z' and 'f a b c', 1-based, 'a/b/c' corners allowed)."""
    verts, tris = [], []
    with open(path, "r", encoding="utf-8") as f:
        for line_no, text in enumerate(f, 1):
            fields = text.split()
            if not fields:
                continue
            if fields[0] == "v":
                if len(fields) < 4:
                    raise ValueError(f"{path}:{line_no}: malformed vertex")
                verts.append((float(fields[1]), float(fields[2]), float(fields[3])))
            elif fields[0] == "f":
                if len(fields) != 4:
                    raise ValueError(f"{path}:{line_no}: face is not a triangle")
                tri = []
                for corner in fields[1:]:
                    idx = int(corner.split("/")[0])
                    if idx < 1 or idx > len(verts):
                        raise ValueError(f"{path}:{line_no}: vertex index out of range")
                    tri.append(idx - 1)
                tris.append(tri)
    return np.asarray(verts, dtype=np.float64).reshape(-1, 3), np.asarray(tris, dtype=np.int64).reshape(-1, 3)


def load_mannequin(body: str = "male", directory: str | None = None) -> Mannequin:
    from .bodies import is_closed_manifold
    d = os.path.abspath(directory or ASSETS)
    obj = os.path.join(d, f"{body}.obj")
    if not os.path.exists(obj):
        raise FileNotFoundError(f"{obj}: cannot open")
    v, t = read_obj_triangles(obj)
    if not is_closed_manifold(t):
        raise ValueError(f"{obj}: mesh is not a closed manifold")
    with open(os.path.join(d, f"{body}.json"), "r", encoding="utf-8") as f:
        doc = json.load(f)
    if doc.get("format") != 1:
        raise ValueError(f"{body}.json: unsupported format {doc.get('format')} (expected 1)")
    for k in GIRTHS:
        if k not in doc["girths"]:
            raise ValueError(f"{body}.json: missing girth '{k}'")
    for k in LANDMARKS:
        if k not in doc["landmarks"]:
            raise ValueError(f"{body}.json: missing landmark '{k}'")
    regions = {}
    for name in G.REGIONS:
        if name not in doc["regions"]:
            raise ValueError(f"{body}.json: missing region '{name}'")
        r = doc["regions"][name]
        regions[name] = G.RegionCylinder(tuple(r["origin"]), tuple(r["axis"]), tuple(r["e0"]), float(r["radius"]))
    return Mannequin(v, t, float(doc["height"]), dict(doc["girths"]),
                     {k: tuple(p) for k, p in doc["landmarks"].items()}, regions)


QUALITY_SOLVER = {"draft": (1, 2), "standard": (1, 2), "fine": (2, 2)}  # substeps, Newton iterations


@dataclasses.dataclass
class TeeScene:
    body: str
    quality: str
    mannequin: Mannequin
    pattern: G.Pattern
    mesh: scenes.Mesh
    scene: scenes.Scene


def tee_scene(body: str = "male", quality: str = "standard", size: str = "M", field=None,
              mannequin_dir: str | None = None) -> TeeScene:
    """The tee on the mannequin, ready to build. `field` reuses an already baked collision field of this body."""
    m = load_mannequin(body, mannequin_dir)
    pattern = G.womens_tee(size) if body == "female" else G.mens_tee(size)
    mesh = G.build_sim_mesh(pattern, G.particle_distance(quality))
    pos = G.place_pattern(pattern, mesh, m.regions)
    fabrics = {"body": scenes.test_jersey(), "rib": test_rib()}
    gar = scenes.Garment(id="tee", mesh=mesh, initial_positions=pos,
                         piece_fabric=[dict(fabrics[p.fabric_slot]) for p in pattern.pieces])
    if field is None:
        from .solver import bake_collision_field
        field = bake_collision_field([(m.vertices, m.triangles)])
    sc = scenes.Scene(garments=[gar], body=field, settings=scenes.settings(), name=f"tee-{body}-{quality}")
    return TeeScene(body, quality, m, pattern, mesh, sc)


def hem_height(ts: TeeScene, x: np.ndarray) -> float:
    """Mean height of the front and back vertices on the pattern's y = 0 edge (tee_drape_helpers.hpp:60-68)."""
    pid = np.asarray([ts.pattern.pieces[int(p)].id for p in ts.mesh.vertex_piece])
    sel = ((pid == "front") | (pid == "back")) & (ts.mesh.rest[:, 1] < 1e-9)
    return float(np.mean(np.asarray(x, dtype=np.float64)[sel, 1]))


def hem_target(ts: TeeScene, size: str = "M") -> float:
    """Mean shoulder-point (HPS) height - body length + take-up, the extra length of the slanted run from the HPS
    over the chest apex (tee_drape_helpers.hpp:40-50)."""
    lm = ts.mannequin.landmarks
    table = G.WOMENS_TEE if ts.body == "female" else G.MENS_TEE
    length = G.resolve_spec(table, size)["body_length"]
    hps, chest = np.asarray(lm["hps_left"]), np.asarray(lm["chest"])
    take_up = float(np.linalg.norm(chest - hps)) - (hps[1] - chest[1])
    return 0.5 * (lm["hps_left"][1] + lm["hps_right"][1]) - length + take_up


@dataclasses.dataclass
class DrapeOptions:
    substeps: int = 1
    iterations: int = 2
    frame_dt: float = 1.0 / 60.0
    max_assembly_seconds: float = 3.0
    seam_gap_done: float = 0.001
    settle_rms_speed: float = 0.002
    settle_hold_frames: int = 60
    max_sim_seconds: float = 20.0
    snapshot_every: int = 30
    max_retries: int = 3
    energy_rise_frames: int = 60


@dataclasses.dataclass
class DrapeResult:
    settled: bool = False
    failed: bool = False
    timed_out: bool = False
    simulated_seconds: float = 0.0
    retries: int = 0
    frames_run: int = 0
    assembly_end_seam_gap: float = 0.0
    assembly_seconds: float = 0.0


ASSEMBLING, SETTLING = 0, 1  # DRAPE_PHASE_ASSEMBLING, DRAPE_PHASE_SETTLING


def options_for(quality: str) -> DrapeOptions:
    s, it = QUALITY_SOLVER[quality.lower()]
    return DrapeOptions(substeps=s, iterations=it)


def drape_to_rest(solver, garments: list, options: DrapeOptions | None = None, on_frame=None) -> DrapeResult:
    """Runs assembling -> settling -> settled (or failed / timed out) on a built solver."""
    o = options or DrapeOptions()
    eps = 1e-9
    res = DrapeResult()
    substeps = o.substeps
    phase = ASSEMBLING
    for g in garments:
        solver.set_phase(g, phase)
    time = phase_time = 0.0
    kept = still = rise = 0
    last_energy = math.inf
    saved = (solver.snapshot(), 0.0, phase, 0.0)
    while True:
        if time >= o.max_sim_seconds - eps:
            res.simulated_seconds, res.timed_out = time, True
            return res
        solver.step(o.frame_dt, substeps, o.iterations)
        res.frames_run += 1
        kept += 1
        time += o.frame_dt
        phase_time += o.frame_dt
        st = solver.stats()
        if on_frame:
            on_frame(res.frames_run, phase, st)
        rise = rise + 1 if st["kinetic_energy"] > last_energy else 0
        last_energy = st["kinetic_energy"]
        if st["has_non_finite"] or rise >= o.energy_rise_frames:
            res.retries += 1
            if res.retries > o.max_retries:
                res.retries = o.max_retries
                res.failed = True
                res.simulated_seconds = time
                return res
            snap, time, phase, phase_time = saved
            solver.restore(snap)  # the snapshot carries the garments' phases and assembly times
            kept = int(time / o.frame_dt + 0.5)
            substeps *= 2
            still = rise = 0
            last_energy = math.inf
            continue
        if phase == ASSEMBLING:
            if st["max_seam_gap"] < o.seam_gap_done or phase_time >= o.max_assembly_seconds - eps:
                res.assembly_end_seam_gap = st["max_seam_gap"]
                res.assembly_seconds = time
                phase = SETTLING
                for g in garments:
                    solver.set_phase(g, phase)
                phase_time = 0.0
                still = rise = 0
                last_energy = math.inf
        else:
            still = still + 1 if st["rms_speed"] < o.settle_rms_speed else 0
            if still >= o.settle_hold_frames:
                res.settled = True
                res.simulated_seconds = time
                return res
        if kept % o.snapshot_every == 0:
            saved = (solver.snapshot(), time, phase, phase_time)
        if time >= o.max_sim_seconds - eps:
            res.simulated_seconds, res.timed_out = time, True
            return res
