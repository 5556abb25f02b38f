"""ctypes mirror of include/drape_hip.h and the loader of the HIP library.

The structs here match the C declarations field for field; `SceneHolder` keeps the numpy arrays a descriptor
points into alive for as long as the descriptor is used.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
PKG_ROOT = os.path.dirname(_HERE)
REPO_ROOT = os.path.dirname(PKG_ROOT)
LIB_PATH = os.path.join(PKG_ROOT, "build", "libdrape_hip.so")

DRAPE_OK = 0
DRAPE_ERR_INVALID_ARGUMENT = 1
DRAPE_ERR_OUT_OF_RANGE = 2
DRAPE_ERR_RUNTIME = 3
DRAPE_ERR_LOGIC = 4

PHASE_ASSEMBLING = 0
PHASE_SETTLING = 1

u32p = C.POINTER(C.c_uint32)
f64p = C.POINTER(C.c_double)
f32p = C.POINTER(C.c_float)
u8p = C.POINTER(C.c_uint8)


class Fabric(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "stretch_warp", "stretch_weft", "stretch_bias", "strain_limit_warp", "strain_limit_weft",
        "bend_warp", "bend_weft", "weight", "thickness", "friction", "damping")]


class Settings(C.Structure):
    _fields_ = [
        ("gravity", C.c_double * 3),
        ("collision_offset", C.c_double),
        ("assembly_closing_speed", C.c_double),
        ("assembly_extra_damping", C.c_double),
        ("damping_scale", C.c_double),
        ("strain_limit_stiffness_scale", C.c_double),
        ("seam_stiffness", C.c_double),
        ("friction_slip_speed", C.c_double),
        ("max_cg_iterations", C.c_int32),
        ("pad0", C.c_int32),
        ("cg_tolerance", C.c_double),
    ]


class MeshDesc(C.Structure):
    _fields_ = [
        ("n_vertices", C.c_uint32),
        ("rest", f64p),
        ("vertex_piece", u32p),
        ("n_triangles", C.c_uint32),
        ("triangles", u32p),
        ("triangle_piece", u32p),
        ("vertex_area", f64p),
        ("boundary", u8p),
        ("n_pieces", C.c_uint32),
        ("piece_grain", f64p),
        ("n_seam_pairs", C.c_uint32),
        ("seam_pairs", u32p),
    ]


class GarmentDesc(C.Structure):
    _fields_ = [
        ("id", C.c_char_p),
        ("mesh", MeshDesc),
        ("initial_positions", f64p),
        ("n_piece_fabrics", C.c_uint32),
        ("piece_fabric", C.POINTER(Fabric)),
    ]


class BakeOptions(C.Structure):
    _fields_ = [("cell_size", C.c_double), ("padding", C.c_double), ("band", C.c_double)]


class BodyMesh(C.Structure):
    _fields_ = [("vertices", C.POINTER(C.c_double)), ("n_vertices", C.c_uint32),
                ("triangles", C.POINTER(C.c_uint32)), ("n_triangles", C.c_uint32)]


class FieldDesc(C.Structure):
    _fields_ = [
        ("origin", C.c_double * 3),
        ("cell_size", C.c_double),
        ("dims", C.c_int32 * 3),
        ("pad0", C.c_int32),
        ("band", C.c_double),
        ("values", f32p),
    ]


class SceneDesc(C.Structure):
    _fields_ = [
        ("n_garments", C.c_uint32),
        ("garments", C.POINTER(GarmentDesc)),
        ("body", C.POINTER(FieldDesc)),
        ("n_pins", C.c_uint32),
        ("pin_particles", u32p),
        ("pin_positions", f64p),
        ("n_forces", C.c_uint32),
        ("force_particles", u32p),
        ("force_vectors", f64p),
        ("settings", Settings),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("has_non_finite", C.c_int32),
        ("pad0", C.c_int32),
        ("kinetic_energy", C.c_double),
        ("rms_speed", C.c_double),
        ("max_seam_gap", C.c_double),
        ("max_body_penetration", C.c_double),
    ]

    def as_dict(self) -> dict:
        return {
            "has_non_finite": bool(self.has_non_finite),
            "kinetic_energy": self.kinetic_energy,
            "rms_speed": self.rms_speed,
            "max_seam_gap": self.max_seam_gap,
            "max_body_penetration": self.max_body_penetration,
        }


class Counters(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in (
        "substeps", "newton_iterations", "pcg_iterations", "line_search_trials",
        "warm_start_fallbacks", "warm_start_accepted")]

    def as_dict(self) -> dict:
        return {n: int(getattr(self, n)) for n, _ in self._fields_}


class PcgReport(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in (
        "solves", "exit_tolerance", "exit_cap", "exit_curvature", "exit_zero", "restarts")] + [
        (n, C.c_double) for n in ("true_rel_max", "true_rel_mean", "true_rel_last", "true_rel_max_converged")] + [
        ("persistent_aborts", C.c_uint64), ("true_rel_excess_max", C.c_double)]

    def as_dict(self) -> dict:
        return {n: (int(v) if isinstance(v, int) else float(v)) for n, _ in self._fields_ for v in [getattr(self, n)]}


class StructureSizes(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in ("n_particles", "n_stretch", "n_hinges", "n_seams", "n_blocks")]


def ptr(a: np.ndarray | None, ctype):
    """Pointer into a contiguous numpy array (None -> NULL)."""
    if a is None:
        return C.cast(None, C.POINTER(ctype))
    assert a.flags["C_CONTIGUOUS"], "array must be C-contiguous"
    return a.ctypes.data_as(C.POINTER(ctype))


def fabric_struct(f: dict) -> Fabric:
    return Fabric(*[float(f[n]) for n, _ in Fabric._fields_])


def settings_struct(s: dict) -> Settings:
    out = Settings()
    out.gravity = (C.c_double * 3)(*s["gravity"])
    for name in ("collision_offset", "assembly_closing_speed", "assembly_extra_damping", "damping_scale",
                 "strain_limit_stiffness_scale", "seam_stiffness", "friction_slip_speed", "cg_tolerance"):
        setattr(out, name, float(s[name]))
    out.max_cg_iterations = int(s["max_cg_iterations"])
    return out


def field_struct(field, keep: list) -> FieldDesc:
    vals = np.ascontiguousarray(field.values, dtype=np.float32).reshape(-1)
    keep.append(vals)
    fd = FieldDesc()
    fd.origin = (C.c_double * 3)(*[float(v) for v in field.origin])
    fd.cell_size = float(field.cell_size)
    fd.dims = (C.c_int32 * 3)(*[int(v) for v in field.dims])
    fd.band = float(field.band)
    fd.values = ptr(vals, C.c_float)
    return fd


class SceneHolder:
    """Builds a SceneDesc from a `scenes.Scene` and owns every buffer it points into."""

    def __init__(self, scene):
        keep: list = []
        self._keep = keep
        garments = (GarmentDesc * max(1, len(scene.garments)))()
        for gi, g in enumerate(scene.garments):
            m = g.mesh
            arrays = {
                "rest": np.ascontiguousarray(m.rest, dtype=np.float64),
                "vertex_piece": np.ascontiguousarray(m.vertex_piece, dtype=np.uint32),
                "triangles": np.ascontiguousarray(m.triangles, dtype=np.uint32),
                "triangle_piece": np.ascontiguousarray(m.triangle_piece, dtype=np.uint32),
                "vertex_area": np.ascontiguousarray(m.vertex_area, dtype=np.float64),
                "boundary": np.ascontiguousarray(m.boundary, dtype=np.uint8),
                "piece_grain": np.ascontiguousarray(m.piece_grain, dtype=np.float64),
                "seam_pairs": np.ascontiguousarray(m.seam_pairs, dtype=np.uint32).reshape(-1, 4),
                "init": np.ascontiguousarray(g.initial_positions, dtype=np.float64),
            }
            # The descriptor carries one count per array kind, so every per-vertex and per-triangle array must
            # have it: a short array would be read past its end (sim_model.cpp:82-85 raises for the positions).
            nv, nt = len(arrays["rest"]), len(arrays["triangles"])
            if arrays["init"].shape != (nv, 3):
                raise ValueError(f"garment '{g.id}': one initial position per mesh vertex required")
            for k, count, per in (("vertex_piece", nv, "vertex"), ("vertex_area", nv, "vertex"),
                                  ("boundary", nv, "vertex"), ("triangle_piece", nt, "triangle")):
                if len(arrays[k]) != count:
                    raise ValueError(f"garment '{g.id}': mesh {k} needs one entry per {per}")
            if arrays["triangles"].ndim != 2 or arrays["triangles"].shape[1] != 3:
                raise ValueError(f"garment '{g.id}': triangles must be (n, 3)")
            keep.append(arrays)
            fabs = (Fabric * max(1, len(g.piece_fabric)))(*[fabric_struct(f) for f in g.piece_fabric])
            keep.append(fabs)
            gid = g.id.encode()
            keep.append(gid)
            md = MeshDesc(
                n_vertices=len(arrays["rest"]),
                rest=ptr(arrays["rest"], C.c_double),
                vertex_piece=ptr(arrays["vertex_piece"], C.c_uint32),
                n_triangles=len(arrays["triangles"]),
                triangles=ptr(arrays["triangles"], C.c_uint32),
                triangle_piece=ptr(arrays["triangle_piece"], C.c_uint32),
                vertex_area=ptr(arrays["vertex_area"], C.c_double),
                boundary=ptr(arrays["boundary"], C.c_uint8),
                n_pieces=len(arrays["piece_grain"]),
                piece_grain=ptr(arrays["piece_grain"], C.c_double),
                n_seam_pairs=len(arrays["seam_pairs"]),
                seam_pairs=ptr(arrays["seam_pairs"], C.c_uint32),
            )
            garments[gi] = GarmentDesc(id=gid, mesh=md, initial_positions=ptr(arrays["init"], C.c_double),
                                       n_piece_fabrics=len(g.piece_fabric), piece_fabric=fabs)
        keep.append(garments)
        desc = SceneDesc()
        desc.n_garments = len(scene.garments)
        desc.garments = garments
        if scene.body is not None:
            fd = field_struct(scene.body, keep)
            keep.append(fd)
            desc.body = C.pointer(fd)
        else:
            desc.body = C.cast(None, C.POINTER(FieldDesc))
        pin_p = np.ascontiguousarray([p for p, _ in scene.pins], dtype=np.uint32)
        pin_x = np.ascontiguousarray([x for _, x in scene.pins], dtype=np.float64).reshape(-1)
        f_p = np.ascontiguousarray([p for p, _ in scene.forces], dtype=np.uint32)
        f_v = np.ascontiguousarray([f for _, f in scene.forces], dtype=np.float64).reshape(-1)
        keep.extend([pin_p, pin_x, f_p, f_v])
        desc.n_pins = len(pin_p)
        desc.pin_particles = ptr(pin_p, C.c_uint32)
        desc.pin_positions = ptr(pin_x, C.c_double)
        desc.n_forces = len(f_p)
        desc.force_particles = ptr(f_p, C.c_uint32)
        desc.force_vectors = ptr(f_v, C.c_double)
        desc.settings = settings_struct(scene.settings)
        self.desc = desc


_lib = None


def load_library(path: str | None = None):
    """Loads libdrape_hip.so; raises RuntimeError naming the path when it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or LIB_PATH
    if not os.path.exists(path):
        raise RuntimeError(f"hip: library not built at {path} (run __graft_entry__.build())")
    lib = C.CDLL(path)
    _declare(lib)
    _lib = lib
    return lib


def _declare(lib):
    S = C.c_void_p
    sig = {
        "drape_last_error": (C.c_char_p, []),
        "drape_settings_default": (C.c_int, [C.POINTER(Settings)]),
        "drape_hip_available": (C.c_int, []),
        "drape_solver_create": (C.c_int, [C.POINTER(S)]),
        "drape_solver_destroy": (None, [S]),
        "drape_solver_build": (C.c_int, [S, C.POINTER(SceneDesc)]),
        "drape_solver_build_batch": (C.c_int, [S, C.c_uint32, C.POINTER(SceneDesc)]),
        "drape_solver_update_fabric": (C.c_int, [S, C.c_char_p, C.c_uint32, C.POINTER(Fabric)]),
        "drape_solver_set_phase": (C.c_int, [S, C.c_char_p, C.c_int32]),
        "drape_solver_range": (C.c_int, [S, C.c_char_p, u32p, u32p]),
        "drape_solver_particle_count": (C.c_int, [S, C.POINTER(C.c_size_t)]),
        "drape_solver_segment_count": (C.c_int, [S, u32p]),
        "drape_solver_segment_range": (C.c_int, [S, C.c_uint32, u32p, u32p]),
        "drape_solver_step": (C.c_int, [S, C.c_double, C.c_int32, C.c_int32]),
        "drape_solver_read_positions": (C.c_int, [S, f32p, C.c_size_t]),
        "drape_solver_read_velocities": (C.c_int, [S, f32p, C.c_size_t]),
        "drape_solver_stats": (C.c_int, [S, C.POINTER(Stats)]),
        "drape_solver_segment_stats": (C.c_int, [S, C.c_uint32, C.c_uint32, C.POINTER(Stats)]),
        "drape_solver_read_positions_normals": (C.c_int, [S, C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_size_t]),
        "drape_bake_field_grid": (C.c_int, [C.POINTER(BodyMesh), C.c_uint32, C.POINTER(BakeOptions),
                                            C.POINTER(FieldDesc)]),
        "drape_bake_collision_field": (C.c_int, [C.POINTER(BodyMesh), C.c_uint32, C.POINTER(BakeOptions),
                                                 C.POINTER(C.c_float)]),
        "drape_vertex_normals": (C.c_int, [C.POINTER(C.c_float), C.c_uint32, C.POINTER(C.c_uint32), C.c_uint32,
                                           C.POINTER(C.c_uint32), C.c_uint32, C.POINTER(C.c_float)]),
        "drape_solver_snapshot_size": (C.c_int, [S, C.POINTER(C.c_size_t)]),
        "drape_solver_snapshot_save": (C.c_int, [S, C.c_void_p, C.c_size_t]),
        "drape_solver_snapshot_restore": (C.c_int, [S, C.c_void_p, C.c_size_t]),
        "drape_solver_counters": (C.c_int, [S, C.POINTER(Counters)]),
        "drape_solver_reset_counters": (C.c_int, [S]),
        "drape_solver_set_graph": (C.c_int, [S, C.c_int32]),
        "drape_solver_set_spmv_sampling": (C.c_int, [S, C.c_int32]),
        "drape_solver_set_pcg_variant": (C.c_int, [S, C.c_int32]),
        "drape_solver_set_pcg_shape": (C.c_int, [S, C.c_int32, C.c_int32, C.c_int32]),
        "drape_solver_set_solve_spin_cap": (C.c_int, [S, C.c_uint32]),
        "drape_solver_persistent_available": (C.c_int, [S, C.POINTER(C.c_int32)]),
        "drape_solver_pcg_shape": (C.c_int, [S, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                             C.POINTER(C.c_int32)]),
        "drape_solver_spmv_timing": (C.c_int, [S, C.POINTER(C.c_uint64), f64p, f64p]),
        "drape_solver_structure_sizes": (C.c_int, [S, C.POINTER(StructureSizes)]),
        "drape_solver_read_structure": (C.c_int, [S, u32p, u32p, u32p, u32p, u32p, u32p]),
        "drape_solver_read_terms": (C.c_int, [S, u32p, u32p, f64p, u32p]),
        "drape_solver_gather_sizes": (C.c_int, [S, u32p, u32p]),
        "drape_solver_pcg_report": (C.c_int, [S, C.POINTER(PcgReport)]),
        "drape_vertex_normals_noalloc": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_void_p,
                                                  C.c_uint32, C.c_void_p, C.c_void_p]),
        "drape_solver_set_refinement": (C.c_int, [S, C.c_int32]),
        "drape_solver_read_gather": (C.c_int, [S, u32p, u32p, u32p, u32p]),
        "drape_solver_debug_assemble": (C.c_int, [S, C.c_double, f32p, f64p, f32p, f32p]),
        "drape_solver_debug_spmv": (C.c_int, [S, f32p, f32p]),
        "drape_solver_debug_system": (C.c_int, [S, f32p, f32p, f32p, u32p, f32p]),
        "drape_net_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_uint64]),
        "drape_net_destroy": (None, [S]),
        "drape_net_param_count": (C.c_int, [S, C.POINTER(C.c_size_t)]),
        "drape_net_get_params": (C.c_int, [S, f32p, C.c_size_t]),
        "drape_net_set_params": (C.c_int, [S, f32p, C.c_size_t]),
        "drape_net_reset_optimizer": (C.c_int, [S]),
        "drape_graph_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_uint32, u32p, u32p, f64p]),
        "drape_graph_destroy": (None, [S]),
        "drape_net_predict": (C.c_int, [S, S, f32p, f32p, f32p, f32p, f32p]),
        "drape_net_train_step": (C.c_int, [S, S, f32p, f32p, f32p, f32p, f32p, C.c_float, f64p, f32p]),
        "drape_solver_set_warm_start": (C.c_int, [S, S]),
        "drape_solver_frame_inputs": (C.c_int, [S, f32p, f32p, f32p, f32p]),
        "drape_solver_mesh_graph_size": (C.c_int, [S, C.POINTER(C.c_size_t)]),
        "drape_solver_mesh_graph": (C.c_int, [S, u32p, u32p]),
        "drape_bsr_spmv": (C.c_int, [C.c_uint32, C.c_uint32, u32p, u32p, f32p, u8p, f32p, f32p, C.c_int32,
                                     f64p]),
        "drape_eval_terms": (C.c_int, [C.c_int32, C.c_uint32, f32p, f32p, C.POINTER(Fabric), C.c_float, f32p,
                                       f32p, f32p]),
        "drape_field_query": (C.c_int, [C.POINTER(FieldDesc), C.c_uint32, f32p, f32p, f32p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args


EXCEPTIONS = {
    DRAPE_ERR_INVALID_ARGUMENT: ValueError,
    DRAPE_ERR_OUT_OF_RANGE: IndexError,
    DRAPE_ERR_RUNTIME: RuntimeError,
    DRAPE_ERR_LOGIC: AssertionError,
}


def check(code: int, last_error) -> None:
    """Maps a C status code to the Python analogue of the reference's exception type."""
    if code == DRAPE_OK:
        return
    msg = last_error()
    if isinstance(msg, bytes):
        msg = msg.decode(errors="replace")
    raise EXCEPTIONS.get(code, RuntimeError)(msg)
